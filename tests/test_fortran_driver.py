"""The Fortran side of the boundary: macaw_gpu.f90 (ISO_C_BINDING) and a
Fortran driver that runs the CMF2 loop on the engine and writes Macaw's
`auto` / `spop` files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from macaw_amd.fortran_io import read_auto, write_descriptor
from macaw_amd.lvc import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "macaw_amd", "fortran")
EXE = os.path.join(ROOT, "macaw_amd", "_build", "fortran", "drive_cmf2")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def build_driver():
    if not os.path.exists(os.path.join(ROOT, "macaw_amd", "libmacaw_gpu.so")):
        from macaw_amd.build import build
        build()
    subprocess.check_call(["make", "-s", "-C", FDIR])
    return EXE


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_binding_builds_and_links():
    exe = build_driver()
    assert os.path.exists(exe)
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libmacaw_gpu.so" in out and "not found" not in out


def test_descriptor_file_layout(tmp_path):
    prob = make_problem("t1", str(tmp_path))
    path = str(tmp_path / "t1.bin")
    write_descriptor(prob, path, 10.0)
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:72], dtype=np.int64)
    assert list(hdr[:8]) == [2, 3, 3, 1, 8, prob.nHam, prob.nAConftot, prob.dimpsi]
    n_i64 = 3 + 3 + 3 + 6 + 3 + 2 * prob.nHam + 3 * prob.nHam
    n_f64 = 8 * 3 + 3 * 8 * 8 * 3 + 2 * prob.nHam + 2 * prob.nAConftot + 2 * prob.dimpsi
    assert len(raw) == 72 + 48 + 8 * n_i64 + 8 * n_f64


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_driver_matches_oracle(tmp_path):
    from oracle_lib import Oracle
    exe = build_driver()
    prob = make_problem("t1", str(tmp_path))
    tf = 3.0 / 0.02418884326505
    write_descriptor(prob, str(tmp_path / "t1.bin"), tf)
    subprocess.check_call([exe, str(tmp_path / "t1.bin")], cwd=str(tmp_path), timeout=300)
    auto = read_auto(str(tmp_path / "auto"))
    outs, pops, _ = Oracle(prob).propagate(0.0, tf)
    assert auto.shape[0] == len(outs)
    for row, o in zip(auto, outs):
        assert abs(row[0] - o.time * 0.02418884326505) < 1e-4
        assert abs(row[1] - o.auto_re) < 1e-8 and abs(row[2] - o.auto_im) < 1e-8
    spop = np.loadtxt(str(tmp_path / "spop"), comments="#")
    assert np.abs(spop[:, 1:] - pops).max() < 2e-7   # f10.7 output
