"""Failure handling of the reference and the engine: the
errors are raised on the host, so every GPU kernel still ends normally.

* missing operator file: input error (input.f90:529-536);
* CMF2 attempts rejected below the minimum step: error instead of the
  reference's endless reject loop (cmf2_base.f90:181-184; the engine's guard);
* Krylov space exhausted (nKrylovMax = 20 reached): error
  (cmf2_integration.f90:64-70);
* a wavefunction that is no longer finite: error at the next output time;
* problems beyond the engine's limits (more than 16 SPFs per mode, more than
  8 electronic states): rejected by mg_create before any device call.
"""
import re

import pytest

from macaw_amd.inputfile import InputError
from macaw_amd.lvc import make_problem
from macaw_amd.problem import load_problem

gpu = pytest.mark.gpu


def test_missing_operator_file(tmp_path):
    make_problem("t1", str(tmp_path))
    inp = (tmp_path / "t1.inp").read_text()
    (tmp_path / "bad.inp").write_text(re.sub(r"Hamiltonian\s*=\s*\S+", "Hamiltonian = nowhere.op", inp))
    with pytest.raises(InputError, match="nowhere.op not found"):
        load_problem(str(tmp_path / "bad.inp"))


@pytest.mark.parametrize("change, message", [
    (dict(nspf=17, npts=20, nmodes=2), "more than 16 SPFs per mode"),
    (dict(nstates=9, gaps=tuple(0.05 * k for k in range(9)), n_lambda=2), "at most 8 electronic states"),
])
def test_unsupported_problems_are_rejected_before_any_device_call(change, message, tmp_path):
    """The engine's size limits are checked on the host before the first HIP
    call, so an unsupported problem gets its own message (here on a machine
    without a GPU, where any device call would fail with a different one)."""
    import dataclasses
    from macaw_amd.engine import Engine, EngineError
    from macaw_amd.lvc import SPECS
    prob = make_problem(dataclasses.replace(SPECS["t1"], **change), str(tmp_path))
    with pytest.raises(EngineError, match=message):
        Engine(prob)


def _bad_lhs(p):
    p.lhs_el = p.lhs_el.copy()
    p.lhs_el[0] = 0


def _bad_rhs(p):
    p.rhs_el = p.rhs_el.copy()
    p.rhs_el[-1] = p.nstates + 1


def _bad_idof(p):
    p.iDOF = p.iDOF.copy()
    p.iDOF[0, 0] = p.ndof + 3


def _zero_spf(p):
    p.nSPF = p.nSPF.copy()
    p.nSPF[1, 0] = 0


def _small_maxdof(p):
    p.nmodeDOF = p.nmodeDOF.copy()
    p.nmodeDOF[0] = p.maxDOF + 1


def _bad_int_type(p):
    p.int_type_psi = 3


def _nan_coefficient(p):
    p.ae = p.ae.copy()
    p.ae[2] = complex("nan")


@pytest.mark.parametrize("mutate, message", [
    (_bad_lhs, "lhs_el / rhs_el of term 1 outside"),
    (_bad_rhs, "lhs_el / rhs_el of term"),
    (_bad_idof, r"iDOF\(1,1\) outside 1..ndof"),
    (_zero_spf, r"nSPF\(2,1\) must be >= 1"),
    (_small_maxdof, r"nmodeDOF\(1\) outside 1..maxdof"),
    (_bad_int_type, r"int_type_psi must be 1 \(Cash-Karp\) or 2 .*integrator not yet implemented for the SPFs"),
    (_nan_coefficient, "coefficient of term 3 is not finite"),
])
def test_malformed_descriptors_are_rejected_before_any_device_call(mutate, message, tmp_path):
    """Out-of-range index tables (states, DOF ids, SPF counts, DOFs per mode)
    and an unknown SPF integrator are refused by mg_create on the host, with
    the offending field named; none of them reaches a kernel's addressing."""
    from macaw_amd.engine import Engine, EngineError
    prob = make_problem("t1", str(tmp_path))
    mutate(prob)
    with pytest.raises(EngineError, match="invalid problem descriptor: " + message):
        Engine(prob)


@gpu
def test_engine_min_step_guard(tmp_path):
    """A rejected attempt whose retry step falls below min_stepsize ends the
    step with an error (the reference's reject loop would shrink h forever).
    The rejection is forced through the attempt replay and min_stepsize is
    set far above any retry step, so the guard fires on the first retry."""
    from macaw_amd.abi import MgAttempt
    from macaw_amd.engine import Engine, EngineError
    prob = make_problem("t1", str(tmp_path))
    prob.min_stepsize = 1.0e6
    e = Engine(prob)
    e.set_replay([MgAttempt(time=0.0, h=1.0, err_A=0.0, err_psi=0.0, accepted=0, nkrylov1=0, nkrylov2=0)])
    with pytest.raises(EngineError, match="minstep"):
        e.cmf2_step(0.0, 1.0)
    # the context stays usable after the error
    out, pops = e.observables(0.0)
    assert abs(sum(pops) - 1.0) < 1e-6


@gpu
def test_engine_krylov_overflow(tmp_path):
    from macaw_amd.engine import Engine, EngineError
    prob = make_problem("t1", str(tmp_path))
    e = Engine(prob)
    with pytest.raises(EngineError, match="Maximum number of Krylov vectors"):
        e.cmf2_step(0.0, 1.0e3)     # ~24 fs in one step: the Lanczos error estimate never falls below 1e-8


@gpu
def test_non_finite_state_is_reported(tmp_path):
    """A NaN in the coefficients is reported at the first output time, before
    any CMF2 step runs on it (the kernels of the observables have no
    data-dependent loops)."""
    import numpy as np
    from macaw_amd.engine import Engine, EngineError
    prob = make_problem("t1", str(tmp_path))
    e = Engine(prob)
    A = prob.A0.copy()
    A[3] = np.nan
    e.set_state(A, prob.psi0)
    with pytest.raises(EngineError, match="not finite"):
        e.propagate(0.0, 1.0 / 0.02418884326505)
