"""bench.py's command-line contract on the CPU: several GPUs are refused with
a message, and the workload line and the contraction flop counts come from
the problem."""
import os
import subprocess
import sys
import tempfile
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_workload_line_and_contraction_flops():
    import bench
    from macaw_amd.lvc import make_problem
    with tempfile.TemporaryDirectory() as d:
        p = make_problem("t2", d)
    line = bench.describe("t2", p)
    assert line.startswith(f"LVC {p.nstates} states x {p.nmodes} modes") and str(p.nAConftot) in line
    calls = []
    eng = types.SimpleNamespace(rho=lambda: calls.append("rho"), meanfield=lambda: calls.append("mf"))
    g = bench.gram_profile(eng, p, 157.3, reps=2)
    # rho: 8 flops per configuration and SPF of every (state, mode)
    expect = sum(8.0 * p.nAConf()[s] * p.nSPF[m, s] for s in range(p.nstates) for m in range(p.nmodes))
    assert g["rho_flops_alg"] == expect
    assert calls.count("rho") == 3 and calls.count("mf") == 3
    assert all(np.isfinite(v) for v in g.values())
