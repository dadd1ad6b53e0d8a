"""Restart files (output.f90:230-241) and INIT-WF ``restart``
(init_calc.f90:290-481): the byte layout read back independently with
struct, round trips, SPF-count remapping, the chosen-state rule, the
normalisation, and the input-file route.  The resumed GPU propagation is in
tests/test_gpu_restart.py.
"""
import dataclasses
import os
import re
import struct

import numpy as np
import pytest

from macaw_amd import restart as R
from macaw_amd.lvc import SPECS, make_problem
from macaw_amd.problem import load_problem


def problem(tmp_path, name="t1", **kw):
    spec = dataclasses.replace(SPECS[name], **kw) if kw else SPECS[name]
    d = tmp_path / f"{spec.name}_{len(kw)}_{kw.get('nspf', '')}"
    d.mkdir(exist_ok=True)
    return make_problem(spec, str(d))


def random_state(prob, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=prob.nAConftot) + 1j * rng.normal(size=prob.nAConftot)
    return A, prob.psi0.copy()


def test_byte_layout(tmp_path):
    prob = problem(tmp_path)
    A, psi = random_state(prob, 1)
    path = tmp_path / "restart"
    R.write_restart(str(path), prob, 41.5, A, psi)
    raw = path.read_bytes()
    pos = 0
    recs = []
    while pos < len(raw):
        (n,) = struct.unpack_from("<i", raw, pos)
        recs.append(raw[pos + 4: pos + 4 + n])
        assert struct.unpack_from("<i", raw, pos + 4 + n)[0] == n
        pos += 8 + n
    ns, nm = prob.nstates, prob.nmodes
    assert len(recs) == 1 + ns + 4
    assert struct.unpack("<2q", recs[0]) == (ns, nm)
    for s in range(ns):
        assert struct.unpack(f"<{nm}q", recs[1 + s]) == tuple(int(v) for v in prob.nSPF[:, s])
    assert struct.unpack("<d", recs[1 + ns]) == (41.5,)
    a = struct.unpack(f"<{2 * A.size}d", recs[2 + ns])
    assert a[0] == A[0].real and a[1] == A[0].imag and a[-1] == A[-1].imag
    assert len(recs[3 + ns]) == 16 * prob.dimpsi
    assert recs[4 + ns] == b""                     # no Gaussian modes


def test_roundtrip_and_identity(tmp_path):
    prob = problem(tmp_path)
    path = str(tmp_path / "restart")
    R.write_restart(path, prob, 3.0, prob.A0, prob.psi0)
    d = R.read_restart(path)
    assert d["time"] == 3.0 and (d["nSPF"] == prob.nSPF).all()
    assert np.array_equal(d["A"], prob.A0) and np.array_equal(d["psi"], prob.psi0)
    A, psi = R.restart_state(prob, d, 0, list(range(1, prob.nstates + 1)))
    assert np.array_equal(psi, prob.psi0)
    assert np.allclose(A, prob.A0, rtol=0, atol=1e-15)


def test_trace_norm_pins(tmp_path):
    prob = problem(tmp_path)
    A, psi = random_state(prob, 2)
    # orthonormal grid SPFs: the norm is |A|^2
    assert abs(R.trace_norm(prob, A, psi) - np.vdot(A, A).real) < 1e-10 * np.vdot(A, A).real
    # Gaussian modes: the single-configuration start is normalised
    g = problem(tmp_path, "g1")
    assert abs(R.trace_norm(g, g.A0, g.psi0) - 1.0) < 1e-12


def test_gaussian_overlap_against_quadrature(tmp_path):
    g = problem(tmp_path, "g1")
    m = next(m for m in range(g.nmodes) if g.gwp_mode(m))
    n = int(g.nSPF[m, 0])
    rng = np.random.default_rng(3)
    xi = (rng.normal(size=(1, n)) + 1j * rng.normal(size=(1, n))) * 0.5
    S = R.gaussian_overlap(g, m, xi)
    x = np.linspace(-40, 40, 200001)
    f = np.exp(np.outer(g.gwp_a[0, :n, m], x ** 2) + np.outer(xi[0], x))
    f = f / np.sqrt(np.trapezoid(np.abs(f) ** 2, x, axis=1))[:, None]
    Sq = np.trapezoid(f.conj()[:, None, :] * f[None, :, :], x, axis=2)
    assert np.abs(S - Sq).max() < 1e-8
    assert np.allclose(np.diag(S), 1.0)


def test_more_spfs_are_orthonormal_and_coefficients_map(tmp_path):
    src = problem(tmp_path)                              # 3 SPFs
    dst = problem(tmp_path, nspf=4)
    A, _ = random_state(src, 4)
    A /= np.sqrt(R.trace_norm(src, A, src.psi0))
    path = str(tmp_path / "restart")
    R.write_restart(path, src, 0.0, A, src.psi0)
    A2, psi2 = R.restart_state(dst, R.read_restart(path), 0, [1, 2])
    ip_s, _ = src.ipsiSPF()
    ip_d, _ = dst.ipsiSPF()
    ng = dst.npar()
    for s in range(dst.nstates):
        for m in range(dst.nmodes):
            n = int(ng[m])
            b_s = src.psi0[ip_s[m, s] - 1: ip_s[m, s] - 1 + 3 * n].reshape(n, 3, order="F")
            b_d = psi2[ip_d[m, s] - 1: ip_d[m, s] - 1 + 4 * n].reshape(n, 4, order="F")
            assert np.array_equal(b_d[:, :3], b_s)
            assert np.allclose(b_d.conj().T @ b_d, np.eye(4), atol=1e-12)
            # the new SPF spans what x-ramp * SPF 3 adds to the first three
            ramp = 0.001 * (np.arange(1, n + 1) - n // 2)
            r = ramp * b_s[:, 2]
            r = r - b_s @ (b_s.conj().T @ r)
            assert abs(abs(np.vdot(b_d[:, 3], r)) - np.linalg.norm(r)) < 1e-10 * np.linalg.norm(r)
    # every source configuration exists in the larger space: A carried over, rest zero
    off_s, off_d = src.AStateOffset(), dst.AStateOffset()
    for s in range(dst.nstates):
        a_d = A2[off_d[s]: off_d[s] + 4 ** dst.nmodes].reshape((4,) * dst.nmodes, order="F")
        a_s = A[off_s[s]: off_s[s] + 3 ** src.nmodes].reshape((3,) * src.nmodes, order="F")
        assert np.allclose(a_d[(slice(0, 3),) * dst.nmodes], a_s, atol=1e-14)
        assert np.abs(a_d).sum() == pytest.approx(np.abs(a_s).sum(), rel=1e-12)
    assert abs(R.trace_norm(dst, A2, psi2) - 1.0) < 1e-12


def test_fewer_spfs_and_chosen_state(tmp_path):
    src = problem(tmp_path)
    dst = problem(tmp_path, nspf=2)
    A, _ = random_state(src, 5)
    path = str(tmp_path / "restart")
    R.write_restart(path, src, 0.0, A, src.psi0)
    A2, psi2 = R.restart_state(dst, R.read_restart(path), 2, [2, 1])   # states swapped, keep state 2 only
    off_s, off_d = src.AStateOffset(), dst.AStateOffset()
    assert np.abs(A2[off_d[0]: off_d[1]]).max() == 0.0
    a_d = A2[off_d[1]:].reshape((2,) * dst.nmodes, order="F")
    a_s = A[off_s[0]: off_s[1]].reshape((3,) * src.nmodes, order="F")[(slice(0, 2),) * src.nmodes]
    assert np.allclose(a_d * np.linalg.norm(a_s), a_s, atol=1e-13)
    assert abs(R.trace_norm(dst, A2, psi2) - 1.0) < 1e-12


def test_mode_count_and_gaussian_checks(tmp_path):
    src = problem(tmp_path)
    path = str(tmp_path / "restart")
    R.write_restart(path, src, 0.0, src.A0, src.psi0)
    d = R.read_restart(path)
    d["nmodes"] = src.nmodes + 1
    with pytest.raises(R.RestartError, match="modes"):
        R.restart_state(src, d, 0, [1, 2])
    with pytest.raises(R.RestartError):
        R.restart_state(src, R.read_restart(path), 0, [1, 3])


def test_input_file_restart_block(tmp_path):
    src = problem(tmp_path)
    A, _ = random_state(src, 6)
    A /= np.sqrt(R.trace_norm(src, A, src.psi0))
    d = tmp_path / "t1_0_"
    R.write_restart(str(d / "restart"), src, 0.0, A, src.psi0)
    inp = (d / "t1.inp").read_text()
    block = "INIT-WF\n   restart\n   restart\n   0\n   1 <-- 1\n   2 <-- 2\nEND-INIT-WF"
    inp2 = re.sub(r"INIT-WF.*?END-INIT-WF", block, inp, flags=re.S)
    (d / "t1r.inp").write_text(inp2)
    p = load_problem(str(d / "t1r.inp"))
    assert np.allclose(p.A0, A, atol=1e-14) and np.array_equal(p.psi0, src.psi0)


def test_gaussian_problem_roundtrip(tmp_path):
    g = problem(tmp_path, "g3")
    path = str(tmp_path / "restart")
    R.write_restart(path, g, 0.0, g.A0, g.psi0)
    d = R.read_restart(path)
    gm = [m for m in range(g.nmodes) if g.gwp_mode(m)]
    ndof = sum(int(g.nmodeDOF[m]) for m in gm)
    ns00 = max(int(g.nSPF[m, :].max()) for m in gm)
    assert d["GWPa"].size == ndof * ns00 * ndof
    widths = g.gwp_a.copy()
    g.gwp_a[:] = 0.0
    A, psi = R.restart_state(g, d, 0, list(range(1, g.nstates + 1)))
    assert np.array_equal(g.gwp_a, widths)
    assert np.array_equal(psi, g.psi0) and np.allclose(A, g.A0, atol=1e-14)


def test_psi_trajectory_records(tmp_path):
    prob = problem(tmp_path)
    path = tmp_path / "psi"
    states = [random_state(prob, 10 + k) for k in range(3)]
    with open(path, "wb") as fh:
        for k, (A, psi) in enumerate(states):
            R.write_psi_record(fh, 20.0 * k, A, psi)
    raw = path.read_bytes()
    n = 8 + 16 * (prob.nAConftot + prob.dimpsi)
    assert len(raw) == 3 * (n + 8) and struct.unpack_from("<i", raw, 0)[0] == n
    recs = R.read_psi_file(str(path), prob.nAConftot, prob.dimpsi)
    assert [r[0] for r in recs] == [0.0, 20.0, 40.0]
    for (t, A, psi), (A0, psi0) in zip(recs, states):
        assert np.array_equal(A, A0) and np.array_equal(psi, psi0)


def test_damaged_files_are_rejected(tmp_path):
    prob = problem(tmp_path)
    path = tmp_path / "restart"
    R.write_restart(str(path), prob, 1.0, prob.A0, prob.psi0)
    raw = path.read_bytes()
    for k, bad in enumerate([raw[:-3],                                  # truncated last marker
                             raw[:len(raw) // 2],                        # truncated inside a record
                             raw[:4] + b"\x00" * 4 + raw[8:],            # damaged first record (nstates = 0)
                             raw[:-4] + struct.pack("<i", 7)]):          # trailing marker disagrees
        p = tmp_path / f"bad{k}"
        p.write_bytes(bad)
        with pytest.raises(R.RestartError):
            R.read_restart(str(p))


def test_c3_jittered_workload(tmp_path):
    """C3 as the workload states it: the generator's jittered Gaussians go
    through a Macaw restart file and come back through INIT-WF restart
    unchanged; the widths are the HO ground state's (A = -1/2), the centres
    q0 = Re xi (for A = -1/2) scatter with the requested deviation 0.2, the
    state is normalised and represents the HO ground state on the top state
    to the accuracy of the jittered basis."""
    from macaw_amd.lvc import make_problem, SPECS
    from macaw_amd.restart import gaussian_overlap, read_restart, trace_norm
    prob = make_problem("c3j", str(tmp_path))
    d = read_restart(str(tmp_path / "c3j.restart"))
    assert np.allclose(d["psi"], prob.psi0) and np.allclose(d["A"], prob.A0, atol=1e-15)
    assert abs(trace_norm(prob, prob.A0, prob.psi0) - 1.0) < 1e-12
    ip, _ = prob.ipsiSPF()
    q0 = []
    for m in range(3):
        assert np.allclose(prob.gwp_a[0, :5, m], -0.5, atol=1e-6)   # width 0.707107 of the input
        for s in range(prob.nstates):
            xi = prob.psi0[int(ip[m, s]) - 1: int(ip[m, s]) - 1 + 5]
            assert np.all(xi.imag == 0.0)
            q0 += list(xi.real / (-2.0 * prob.gwp_a[0, :5, m]))
    q0 = np.array(q0)
    assert len(q0) == 45 and 0.13 < q0.std() < 0.27 and abs(q0.mean()) < 0.1
    # only the top state is populated, as a product of per-mode expansions
    off = prob.AStateOffset()
    top = prob.nstates - 1
    assert np.abs(prob.A0[: off[top]]).max() == 0.0
    # fidelity with the HO ground state: per Gaussian mode |<g0|P g0>| with P
    # the projector on the jittered Gaussians
    for m in range(3):
        xi = prob.psi0[int(ip[m, top]) - 1: int(ip[m, top]) - 1 + 5].reshape(1, 5)
        xa = np.concatenate([xi, np.zeros((1, 1), complex)], axis=1)
        from macaw_amd.lvc import _overlap_with_ground
        S6 = _overlap_with_ground(prob, m, xa)
        c = np.linalg.solve(S6[:5, :5], S6[:5, 5])
        fid = (S6[5, :5] @ c).real
        assert 1.0 - 1e-6 < fid <= 1.0 + 1e-9
    assert SPECS["c3"].eps_C == 1e-8


def test_c3_jittered_spf_error_floor(tmp_path):
    """Why the jittered C3 cannot run at eps_int = 1e-7: the SPF error
    estimate (DeltaPsi of the two half-step SPF sets, cmf2_base.f90:140)
    stops falling with the step at its rounding floor of a few 1e-8, while
    the A error falls as h^2; at h = 1e-3 fs it is already above 1e-7."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__))))
    from oracle_lib import Oracle
    from macaw_amd.lvc import make_problem
    prob = make_problem("c3j", str(tmp_path))
    prob.eps_int = 1e6
    errs = {}
    for h_fs in (1e-3, 1e-5, 1e-6):
        rc, hd, hn, atts = Oracle(prob).cmf2_step(0.0, h_fs / 0.02418884326505)
        assert rc == 0
        errs[h_fs] = (atts[0].err_A, atts[0].err_psi)
    assert errs[1e-3][1] > 1e-7
    assert 5e-9 < errs[1e-6][1] < 1e-7 and 5e-9 < errs[1e-5][1] < 1e-7
    assert errs[1e-6][0] < 1e-12
