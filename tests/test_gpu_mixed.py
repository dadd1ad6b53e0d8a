"""Mixed-precision form (precision 1) of the HIP engine vs the FP64 oracle.

The mixed form keeps the A-vector, SPFs, density and mean-field matrices and
every reduction in FP64 and runs the H.A chains, the Krylov vectors beyond the
first, the mean-field mode products and the density-matrix Grams on complex
FP32 data.  Bars:
  * kernels: norm-wise relative error <= 1e-5 (FP32 rounding over a chain of
    up to 8 mode products; a wrong term or index is an O(1) error);
  * h1mode (FP64 SPFs): 1e-12, as in the FP64 engine; rho (Gram of the
    FP32-rounded A, accumulated in FP64): 1e-6 norm-wise;
  * trajectories: max |dC(t)|, |dpop|, |dnorm| <= 1e-6 against the FP64
    oracle, free-running and with the oracle's attempt sequence replayed.
"""
import numpy as np
import pytest

from oracle_lib import Oracle
from macaw_amd.lvc import make_problem

gpu = pytest.mark.gpu
CASES = ["t1", "t2", "t3", "t4", "t5"]
AU2FS = 0.02418884326505


def mixed_engine(prob):
    from macaw_amd.engine import Engine
    return Engine(prob, precision=1)


def rand_c(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def perturbed_state(prob, seed):
    rng = np.random.default_rng(seed)
    A = rand_c(rng, prob.nAConftot)
    A /= np.linalg.norm(A)
    psi = prob.psi0 + 0.05 * rand_c(rng, prob.dimpsi)
    return A, psi


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.fixture(scope="module", params=CASES)
def case(request, tmp_path_factory):
    return make_problem(request.param, str(tmp_path_factory.mktemp(request.param)))


@gpu
def test_mixed_matrices_match_oracle(case):
    A, psi = perturbed_state(case, 1)
    o = Oracle(case)
    o.set_state(A, psi)
    e = mixed_engine(case)
    e.set_state(A, psi)
    h_o = o.h1mode()
    h_e = e.h1mode()
    assert np.abs(h_e - h_o).max() <= 1e-12 * max(1.0, np.abs(h_o).max())
    r_o = o.rho()
    r_e = e.rho()
    assert rel(r_e, r_o) <= 1e-6          # FP32-rounded A, FP64 accumulation
    mf_o, rmf_o, _ = o.meanfield()
    mf_e, rmf_e = e.meanfield()
    assert rel(mf_e, mf_o) <= 1e-5
    assert rel(rmf_e, rmf_o) <= 1e-4


@gpu
def test_mixed_apply_h_matches_oracle(case):
    A, psi = perturbed_state(case, 2)
    o = Oracle(case)
    o.set_state(A, psi)
    o.h1mode()
    e = mixed_engine(case)
    e.set_state(A, psi)
    e.h1mode()
    rng = np.random.default_rng(3)
    for _ in range(3):
        B = rand_c(rng, case.nAConftot)
        err = rel(e.apply_h(B), o.apply_h(B))
        assert err <= 1e-5
        assert err > 0.0          # the FP32 path really ran (FP64 would agree to ~1e-15)


def max_errors(outs_e, pops_e, outs_o, pops_o):
    assert len(outs_e) == len(outs_o)
    dc = max(abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im)) for a, b in zip(outs_e, outs_o))
    dn = max(abs(a.norm - b.norm) for a, b in zip(outs_e, outs_o))
    dp = float(np.abs(pops_e - pops_o).max())
    return dc, dn, dp


@gpu
@pytest.mark.parametrize("name,tf_fs", [("t1", 4.0), ("t2", 2.0), ("t3", 2.0), ("t4", 3.0), ("t5", 1.0), ("c1", 2.0)])
def test_mixed_propagation_within_1e6_of_fp64(name, tf_fs, tmp_path):
    prob = make_problem(name, str(tmp_path))
    tf = tf_fs / AU2FS
    o = Oracle(prob)
    outs_o, pops_o, atts_o = o.propagate(0.0, tf)
    e = mixed_engine(prob)
    outs_e, pops_e, atts_e = e.propagate(0.0, tf)
    dc, dn, dp = max_errors(outs_e, pops_e, outs_o, pops_o)
    print(f"{name}: free-running max|dC|={dc:.2e} |dnorm|={dn:.2e} |dpop|={dp:.2e} "
          f"attempts {len(atts_e)} vs {len(atts_o)}")
    assert dc <= 1e-6 and dn <= 1e-6 and dp <= 1e-6
    assert abs(len(atts_e) - len(atts_o)) <= 0.25 * len(atts_o) + 2
    e2 = mixed_engine(prob)
    e2.set_replay(atts_o)
    outs_r, pops_r, _ = e2.propagate(0.0, tf)
    dc, dn, dp = max_errors(outs_r, pops_r, outs_o, pops_o)
    print(f"{name}: replayed max|dC|={dc:.2e} |dnorm|={dn:.2e} |dpop|={dp:.2e}")
    assert dc <= 1e-6 and dn <= 1e-6 and dp <= 1e-6


@gpu
def test_mixed_full_size_apply_h_against_fp64_engine():
    """At the 4-state 8-mode 8^8 shape the mixed H.A agrees with the FP64
    engine's to FP32 accuracy, and stays Hermitian to that accuracy."""
    import dataclasses
    import tempfile
    from macaw_amd.engine import Engine
    from macaw_amd.lvc import SPECS
    spec = dataclasses.replace(SPECS["c2"], name="c2m")
    with tempfile.TemporaryDirectory() as d:
        prob = make_problem(spec, d)
    rng = np.random.default_rng(11)
    n = prob.nAConftot
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(n)
    y = (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(n)
    e64 = Engine(prob)
    e64.h1mode()
    hx64 = e64.apply_h(x)
    e64.close()
    e = mixed_engine(prob)
    e.h1mode()
    hx = e.apply_h(x)
    hy = e.apply_h(y)
    assert rel(hx, hx64) <= 1e-5
    lhs, rhs = np.vdot(x, hy), np.vdot(hx, y)
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1e-3)
