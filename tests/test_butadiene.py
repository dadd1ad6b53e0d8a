"""Macaw's own Butadiene example (Examples/Butadiene_5modes_CMF and
Examples/Butadiene_5modes, copied as data into tests/golden/butadiene_5modes
by tests/golden/copy_butadiene.py): 2 states, 5 DOFs (a sine DVR, an HO DVR
and three Gaussian DOFs in 3 modes, one of them a 2-D Gaussian mode), the
Hamiltonian with tabulated-and-splined operators (`interp`), three
expectation operators; CMF2 in one deck, Dormand-Prince in the other.

Pins (oracle, CPU): the two decks' integrators, which share nothing but the
Hamiltonian, agree on C(t) to their tolerances; CMF2 keeps the norm and the
energy; <Q5^2> of the initial Gaussian of width 1/sqrt(2) on the sine DVR is
1/2 and <H> from the Hamiltonian read as an expectation operator is the
output record's energy.  The reference holds no output files for the
example, so the trajectories themselves are checked engine against oracle
(GPU).
"""
import os

import numpy as np
import pytest

from macaw_amd.problem import load_problem, with_operator

gpu = pytest.mark.gpu
AU2FS = 0.02418884326505
HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "butadiene_5modes")
REF = "/root/reference/Examples"


def deck(name="but_5D.inp"):
    return load_problem(os.path.join(DECKS, name))


def _auto(outs):
    return np.array([complex(o.auto_re, o.auto_im) for o in outs])


def _expect(o, prob, A, psi, terms):
    """<Psi|O|Psi> through the oracle's H.A with O, in the basis the
    Gaussian modes' D0 orthonormalises (B = D^H S A)."""
    from oracle_lib import Oracle
    oo = Oracle(with_operator(prob, terms))
    oo.set_state(A, psi)
    oo.prepare_gwp()
    _, D = oo.gwp_overlaps()
    B = oo.a_to_b(A, D)
    return complex(np.vdot(B, oo.apply_h(B)))


@pytest.mark.skipif(not os.path.isdir(REF), reason="Macaw examples not present")
def test_fixture_is_the_reference_example():
    for f in sorted(os.listdir(DECKS)):
        src = os.path.join(REF, "Butadiene_5modes" if f == "but_5D_dopri.inp" else "Butadiene_5modes_CMF",
                           "but_5D.inp" if f == "but_5D_dopri.inp" else f)
        assert open(src, "rb").read() == open(os.path.join(DECKS, f), "rb").read(), f


def test_decks_load():
    p, q = deck(), deck("but_5D_dopri.inp")
    for prob in (p, q):
        assert prob.nstates == 2 and prob.ndof == 5 and prob.nmodes == 3
        assert list(prob.ngp) == [56, 22, 20, 25, 20] and list(prob.dof_gwp) == [0, 1, 1, 0, 1]
        assert prob.nAConftot == 6 * 6 * 7 + 5 * 6 * 7
        # q2.interp and qcoup.interp, both on the sine-DVR DOF
        assert prob.gridoper.shape == (56, 2, 5)
        assert [f for f, _ in prob.expect] == ["e1.op", "e12.op", "q1.op"]
    assert p.int_type == 10 and q.int_type == 2
    # the tabulated parabola of q2.interp on the sine grid, inside the table
    x = p.grid[:56, 0]
    inside = np.abs(x) <= 7.0
    assert np.abs(p.gridoper[:56, 0, 0][inside] - x[inside] ** 2).max() < 0.05


def test_oracle_cmf2_and_dopri_agree():
    from oracle_lib import Oracle
    tf = 1.0 / AU2FS
    p, q = deck(), deck("but_5D_dopri.inp")
    oc, _, ac = Oracle(p).propagate(0.0, tf)
    od, _, ad = Oracle(q).propagate_vmf(2, 0.0, tf)
    cc, cd = _auto(oc), _auto(od)
    assert len(cc) == len(cd) == 3
    assert np.abs(cc - cd).max() < 1e-4
    # the free motion is not trivial over 1 fs
    assert abs(cc[-1] - 1.0) > 0.5
    assert np.allclose([o.norm for o in oc], 1.0, atol=1e-9)
    e = np.array([o.energy for o in oc]) * 27.2113961
    assert np.abs(e - e[0]).max() < 1e-6
    assert all(a.accepted in (0, 1) for a in ac) and sum(a.accepted for a in ac) >= 2


def test_oracle_initial_expectation_values():
    from oracle_lib import Oracle
    p = deck()
    o = Oracle(p)
    out, _ = o.observables(0.0)
    A, psi = o.get_state()
    vals = {f: _expect(o, p, A, psi, terms) for f, terms in p.expect}
    # <Q5^2> of exp(-x^2) (width 1/sqrt(2) in the amplitude's exponent)
    assert abs(vals["q1.op"] - 0.5) < 1e-6
    # H as an expectation operator: the energy of the first output record
    h = _expect(o, p, A, psi, _ham_terms(p))
    assert abs(h.real - out.energy * out.norm) < 1e-10
    assert all(abs(v.imag) < 1e-12 for v in vals.values())


def _ham_terms(prob):
    from macaw_amd.inputfile import OpParams, parse_operator_file
    from macaw_amd.problem import _find_file
    return parse_operator_file(_find_file(os.path.join(DECKS, "Ham_interp.op")), prob.nstates, OpParams())


@gpu
def test_engine_cmf2_matches_oracle():
    """2 fs of the CMF2 deck, the oracle's attempts replayed: C(t),
    populations, energies and the end state."""
    from oracle_lib import Oracle
    from macaw_amd.engine import Engine
    p = deck()
    tf = 2.0 / AU2FS
    o = Oracle(p)
    oo, po, oa = o.propagate(0.0, tf)
    e = Engine(p)
    e.set_replay(oa)
    eo, pe, _ = e.propagate(0.0, tf)
    assert len(eo) == len(oo) == 5
    assert np.abs(_auto(eo) - _auto(oo)).max() < 1e-10
    assert np.abs(np.asarray(pe) - np.asarray(po)).max() < 1e-10
    assert max(abs(a.energy - b.energy) for a, b in zip(eo, oo)) < 1e-12
    e.close()


@gpu
def test_engine_cmf2_free_run_matches_oracle():
    """The engine's own step controller (classical, as the reference) over
    1 fs: the same attempts and C(t) to the rounding of the error tests."""
    from oracle_lib import Oracle
    from macaw_amd.engine import Engine
    p = deck()
    tf = 1.0 / AU2FS
    oo, _, oa = Oracle(p).propagate(0.0, tf)
    e = Engine(p)
    eo, _, ea = e.propagate(0.0, tf)
    assert [a.accepted for a in ea] == [a.accepted for a in oa]
    assert max(abs(a.h - b.h) / b.h for a, b in zip(ea, oa)) < 1e-6
    assert np.abs(_auto(eo) - _auto(oo)).max() < 1e-8
    e.close()


@gpu
def test_engine_dopri_matches_oracle():
    from oracle_lib import Oracle
    from macaw_amd.engine import Engine
    q = deck("but_5D_dopri.inp")
    tf = 1.0 / AU2FS
    oo, po, oa = Oracle(q).propagate_vmf(2, 0.0, tf)
    e = Engine(q)
    e.set_replay(oa)
    eo, pe, _ = e.propagate_vmf(2, 0.0, tf)
    assert len(eo) == len(oo)
    assert np.abs(_auto(eo) - _auto(oo)).max() < 1e-9
    assert np.abs(np.asarray(pe) - np.asarray(po)).max() < 1e-9
    e.close()


@gpu
def test_engine_expectation_values_match_oracle():
    from oracle_lib import Oracle
    from macaw_amd import run as RUN
    from macaw_amd.engine import Engine
    p = deck()
    e = Engine(p)
    e.propagate(0.0, 1.0 / AU2FS)
    A, psi = e.get_state()
    vals = RUN.Expectations(p)(A, psi)
    o = Oracle(p)
    for (f, terms), v in zip(p.expect, vals):
        ref = _expect(o, p, A, psi, terms)
        assert abs(v - ref) < 1e-10 * max(1.0, abs(ref)), (f, v, ref)
    e.close()
