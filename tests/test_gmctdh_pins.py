"""Pins of the oracle's G-MCTDH (Gaussian-SPF) restatement.

Every check here is independent of the oracle's own formulas: Gaussian
integrals against trapezoidal quadrature of the actual functions, the
regularised solve and the Cholesky inverse square root against numpy,
A_to_B / B_to_A against their defining identity, and whole propagations
against closed forms (a displaced harmonic oscillator's coherent state, which
frozen-width Gaussians of the ground-state width represent exactly) and a
separable product (Gaussian mode x grid mode) whose grid factor is
propagated exactly with expm.
"""
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle_lib import Oracle, gau_int, gau_ovl, inverse_sqrt_u, std_reg_sol_tikh
from macaw_amd.lvc import make_problem
from macaw_amd.problem import load_problem

AU2EV = 27.2113961
AU2FS = 0.02418884326505


def _grid(A1, A2, xi1, xi2, half=40.0, n=40001):
    s = A1 + A2
    xc = -(np.conj(xi1) + xi2).real / (2 * s)
    sig = 1.0 / np.sqrt(-2 * s)
    x = np.linspace(xc - half * sig, xc + half * sig, n)
    return x, x[1] - x[0]


def _quad_ratio(A1, A2, xi1, xi2, op, deriv=False):
    """<g1|O|g2>/<g1|g2> (deriv: <x g1|O|g2>/<g1|g2>) by the trapezoidal rule."""
    x, h = _grid(A1, A2, xi1, xi2)
    g1 = np.exp(A1 * x * x + xi1 * x)
    g2 = np.exp(A2 * x * x + xi2 * x)
    d1 = (2 * A2 * x + xi2) * g2
    d2 = ((2 * A2 * x + xi2) ** 2 + 2 * A2) * g2
    if op == 0:
        og2 = g2
    elif op == -1:
        og2 = -1j * d1
    elif op == -2:
        og2 = d2
    elif op == -3:
        og2 = -1j * (2 * x * d1 + g2)
    else:
        og2 = x ** op * g2
    bra = np.conj(g1) * (x if deriv else 1.0)
    return np.sum(bra * og2) * h / (np.sum(np.conj(g1) * g2) * h)


CASES = [(-0.5, -0.5, 0.3 + 0.2j, -0.4 + 0.1j), (-0.3, -0.8, 1.1 - 0.5j, 0.2 + 0.7j),
         (-1.2, -0.25, -0.6 + 1.5j, 0.9 - 0.3j)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("op", [0, -1, -2, -3] + list(range(1, 21)))
def test_gau_int_matches_quadrature(case, op):
    A1, A2, x1, x2 = case
    ref = _quad_ratio(A1, A2, x1, x2, op)
    got = gau_int(A1, A2, x1, x2, op)
    assert abs(got - ref) <= 1e-10 * max(1.0, abs(ref))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("op", [0, -1, -2, -3] + list(range(1, 20)))
def test_gau10_int_matches_quadrature(case, op):
    A1, A2, x1, x2 = case
    ref = _quad_ratio(A1, A2, x1, x2, op, deriv=True)
    got = gau_int(A1, A2, x1, x2, op, deriv=True)
    assert abs(got - ref) <= 1e-10 * max(1.0, abs(ref))


def test_gau_int_rejects_parametrised_operators():
    with pytest.raises(ValueError):
        gau_int(-0.5, -0.5, 0.1, 0.2, 201)


@pytest.mark.parametrize("equal", [True, False])
def test_gau_ovl_matches_quadrature(equal):
    a1 = np.array([-0.5, -0.7])
    a2 = a1.copy() if equal else np.array([-0.3, -1.1])
    x1 = np.array([0.4 + 0.3j, -0.2 + 0.9j])
    x2 = np.array([-0.5 - 0.2j, 0.6 + 0.1j])
    ref = 1.0 + 0j
    for k in range(2):
        x, h = _grid(a1[k], a2[k], x1[k], x2[k])
        g1 = np.exp(a1[k] * x * x + x1[k] * x)
        g2 = np.exp(a2[k] * x * x + x2[k] * x)
        n1 = np.sqrt(np.sum(np.abs(g1) ** 2) * h)
        n2 = np.sqrt(np.sum(np.abs(g2) ** 2) * h)
        ref *= np.sum(np.conj(g1) * g2) * h / (n1 * n2)
    assert abs(gau_ovl(a1, a2, x1, x2) - ref) < 1e-12


def test_std_reg_sol_tikh_matches_eigh():
    rng = np.random.default_rng(5)
    M = rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6))
    C = M @ M.conj().T
    C[:, 0] *= 1e-9
    C[0, :] *= 1e-9                      # nearly singular direction
    y = rng.normal(size=6) + 1j * rng.normal(size=6)
    eps = 1e-6
    w, U = np.linalg.eigh(C)
    ref = U @ ((w / (w * w + eps * eps)) * (U.conj().T @ y))
    # only the upper triangle is read
    Cu = np.triu(C) + np.tril(np.full((6, 6), 7.0 + 7j), -1)
    assert np.abs(std_reg_sol_tikh(Cu, y, eps) - ref).max() < 1e-8 * np.abs(ref).max()


def test_inverse_sqrt_u_is_cholesky_inverse():
    rng = np.random.default_rng(6)
    M = rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5))
    S = M @ M.conj().T + 5 * np.eye(5)
    D = inverse_sqrt_u(S)
    assert np.allclose(np.tril(D, -1), 0)
    assert np.abs(D.conj().T @ S @ D - np.eye(5)).max() < 1e-12
    assert np.abs(D @ D.conj().T - np.linalg.inv(S)).max() < 1e-12


def test_overlaps_and_a_b_transforms(tmp_path):
    prob = make_problem("g3", str(tmp_path))
    o = Oracle(prob)
    S, D = o.gwp_overlaps()
    rng = np.random.default_rng(7)
    A = rng.normal(size=prob.nAConftot) + 1j * rng.normal(size=prob.nAConftot)
    for m in range(2):                   # the Gaussian modes
        for s in range(prob.nstates):
            n = prob.nSPF[m, s]
            Sm = S[s, s, m, :n, :n].T
            Dm = D[s, m, :n, :n].T
            assert np.abs(np.diag(Sm) - 1).max() == 0
            assert np.abs(Dm.conj().T @ Sm @ Dm - np.eye(n)).max() < 1e-12
    # B = D^H S A, A = D B: the round trip is the identity (D D^H = S^-1)
    B = o.a_to_b(A, D)
    assert np.abs(o.b_to_a(B, D) - A).max() < 1e-11 * np.abs(A).max()
    # <Psi|Psi> = A^H (x S) A = |B|^2
    assert abs(o.norm_wf(A, prob.psi0) - np.vdot(B, B).real) < 1e-10 * np.vdot(B, B).real


def test_err_spf_of_a_gaussian_mode_is_the_norm_of_the_change(tmp_path):
    """Calc_Err_SPF of the Gaussian mode of g1 (the other modes are grid
    modes with orthonormal SPFs): || Psi(conj A, xi1) - Psi(conj A, xi2) ||
    (the reference's trace pairs rho with the transposed overlaps, as on grid
    modes), with the normalised Gaussians exp(a x^2 + xi x) sampled on a fine
    grid."""
    import brute
    prob = make_problem("g1", str(tmp_path))
    assert prob.gwp_mode(0) and int(prob.nmodeDOF[0]) == 1
    rng = np.random.default_rng(8)
    A = rng.normal(size=prob.nAConftot) + 1j * rng.normal(size=prob.nAConftot)
    A /= np.linalg.norm(A)
    ip, _ = prob.ipsiSPF()
    nsets = prob.nstates if prob.MultiSet[0] else 1
    psi1, psi2 = prob.psi0, prob.psi0.copy()
    for s in range(nsets):
        n = int(prob.nSPF[0, s])
        psi2[ip[0, s] - 1: ip[0, s] - 1 + n] += 0.05 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    o = Oracle(prob)
    o.set_state(A, psi1)
    o.prepare_gwp()
    o.rho()
    err = o.err_spf(0, psi1, psi2)

    x = np.linspace(-15.0, 15.0, 30001)
    h = x[1] - x[0]

    def gaussians(psi, s):
        jel = s if prob.MultiSet[0] else 0
        n = int(prob.nSPF[0, jel])
        a = prob.gwp_a[0, :n, 0]
        xi = psi[ip[0, jel] - 1: ip[0, jel] - 1 + n]
        assert np.abs(-xi.real / (2 * a)).max() < 5.0          # centres well inside the grid
        G = np.exp(np.outer(x * x, a) + np.outer(x, xi))
        return G / np.sqrt(np.sum(np.abs(G) ** 2, axis=0) * h)

    def change_norm(Av):
        total = 0.0
        for s in range(prob.nstates):
            T = brute.state_tensor(prob, Av, s)
            T = brute.apply_mode(T, gaussians(psi1, s) - gaussians(psi2, s), 0)
            for m in range(1, prob.nmodes):
                T = brute.apply_mode(T, brute.spf_block(prob, psi1, m, s), m)
            total += np.sum(np.abs(T) ** 2) * h
        return np.sqrt(total)

    assert err > 1e-2
    assert err == pytest.approx(change_norm(A.conj()), rel=1e-9)
    assert abs(err - change_norm(A)) > 1e-3 * err


def _coherent_problem(d, ng, w_ev=0.2, k_ev=0.1):
    with open(os.path.join(d, "h.op"), "w") as fh:
        fh.write(f"{-w_ev / 2 / AU2EV:.15e} | 1 |1 dq^2\n{w_ev / 2 / AU2EV:.15e} | 1 |1 q^2\n"
                 f"{k_ev / AU2EV:.15e} | 1 |1 q\n----\n")
    with open(os.path.join(d, "h.inp"), "w") as fh:
        fh.write(f"""tfinal = 5.0
tout = 0.5
nstates = 1
ndof = 1
nmodes = 1
SPF
 M, 1 = {ng}
GRID
  1  gwp  41  -8.0  8.0
Hamiltonian = h.op
INIT-WF
   single-conf
   el  1
   1   gwp 0  0.0  0.0  0.707107   ovl = 0.7
END-INIT-WF
eps_int = 1.0d-7
eps_C = 1.0d-8
integration = CMF2
Lanczos  1.d-7
rk45     1.0d-7
""")
    return load_problem(os.path.join(d, "h.inp"))


@pytest.mark.parametrize("ng", [2, 3])
def test_coherent_state_is_exact(ng, tmp_path):
    """H = w/2 (-d^2 + q^2) + k q from the ground state: C(t) =
    exp(-b^2 (1 - e^{-iwt})) e^{-i(w/2 - k^2/2w) t}, b = k / (w sqrt 2).
    The populated Gaussian keeps the ground-state width, so G-MCTDH is
    exact; the other Gaussians are unpopulated (regularised C matrix)."""
    prob = _coherent_problem(str(tmp_path), ng)
    w, k = 0.2 / AU2EV, 0.1 / AU2EV
    outs, _, _ = Oracle(prob).propagate(0.0, 5.0 / AU2FS)
    b2 = (k / (w * np.sqrt(2))) ** 2
    for x in outs:
        t = x.time
        c = np.exp(-b2 * (1 - np.exp(-1j * w * t))) * np.exp(-1j * (w / 2 - k * k / (2 * w)) * t)
        assert abs(c - complex(x.auto_re, x.auto_im)) < 2e-7
        assert abs(x.norm - 1.0) < 1e-9
        assert abs(x.energy - outs[0].energy) < 1e-9


def test_separable_product_matches_exact(tmp_path):
    """Gaussian mode (3 Gaussians, displaced oscillator) x grid mode (HO-DVR
    with a q^2 and q perturbation), no coupling: MCTDH is exact for a
    product state, so C(t) = C_gauss(t) C_grid(t), the grid factor from expm
    of the one-dimensional DVR Hamiltonian."""
    d = str(tmp_path)
    w1, k1, w2, k2, g2 = 0.2, 0.1, 0.15, -0.08, 0.02
    with open(os.path.join(d, "p.op"), "w") as fh:
        fh.write(f"{-w1 / 2 / AU2EV:.15e} | 1 |1 dq^2\n{w1 / 2 / AU2EV:.15e} | 1 |1 q^2\n"
                 f"{k1 / AU2EV:.15e} | 1 |1 q\n"
                 f"{-w2 / 2 / AU2EV:.15e} | 1 |2 dq^2\n{(w2 / 2 + g2) / AU2EV:.15e} | 1 |2 q^2\n"
                 f"{k2 / AU2EV:.15e} | 1 |2 q\n----\n")
    with open(os.path.join(d, "p.inp"), "w") as fh:
        fh.write("""tfinal = 4.0
tout = 0.5
nstates = 1
ndof = 2
nmodes = 2
SPF
 M, 1 = 3
 M, 2 = 3
GRID
  1  gwp  41  -8.0  8.0
  2  HO  24  0.0  1.0
Hamiltonian = p.op
INIT-WF
   single-conf
   el  1
   1   gwp 0  0.0  0.0  0.707107   ovl = 0.7
   2   gauss  0.0  0.0  0.707107
END-INIT-WF
eps_int = 1.0d-7
eps_C = 1.0d-8
integration = CMF2
Lanczos  1.d-7
rk45     1.0d-7
""")
    prob = load_problem(os.path.join(d, "p.inp"))
    outs, _, _ = Oracle(prob).propagate(0.0, 4.0 / AU2FS)
    W1, K1 = w1 / AU2EV, k1 / AU2EV
    b2 = (K1 / (W1 * np.sqrt(2))) ** 2
    n = 24
    x = prob.grid[:n, 1]
    H2 = -(w2 / 2 / AU2EV) * prob.d2[:n, :n, 1] + np.diag((w2 / 2 + g2) / AU2EV * x * x + k2 / AU2EV * x)
    ip, _ = prob.ipsiSPF()
    phi0 = prob.psi0[ip[1, 0] - 1: ip[1, 0] - 1 + n]
    for o in outs:
        t = o.time
        c1 = np.exp(-b2 * (1 - np.exp(-1j * W1 * t))) * np.exp(-1j * (W1 / 2 - K1 * K1 / (2 * W1)) * t)
        c2 = np.vdot(phi0, expm(-1j * H2 * t) @ phi0) / np.vdot(phi0, phi0)
        assert abs(c1 * c2 - complex(o.auto_re, o.auto_im)) < 5e-7


@pytest.mark.parametrize("name", ["g1", "g2"])
def test_gmctdh_conserves_norm_and_energy(name, tmp_path):
    prob = make_problem(name, str(tmp_path))
    outs, pops, _ = Oracle(prob).propagate(0.0, 2.0 / AU2FS)
    for o, p in zip(outs, pops):
        assert abs(o.norm - 1.0) < 1e-8
        assert abs(p.sum() - o.norm) < 1e-12
        assert abs(o.energy - outs[0].energy) < 1e-8
    assert pops[-1][0] > 1e-4            # population moves through the coupling


@pytest.mark.parametrize("method", [1, 2])
def test_vmf_coherent_state_is_exact(method, tmp_path):
    """VMF G-MCTDH (derivatives.f90 Gaussian branch: S^-1 H A - tau A and the
    C/Y parameter equations) on the displaced-oscillator coherent state."""
    prob = _coherent_problem(str(tmp_path), 3)
    w, k = 0.2 / AU2EV, 0.1 / AU2EV
    outs, _, _ = Oracle(prob).propagate_vmf(method, 0.0, 3.0 / AU2FS)
    b2 = (k / (w * np.sqrt(2))) ** 2
    for x in outs:
        t = x.time
        c = np.exp(-b2 * (1 - np.exp(-1j * w * t))) * np.exp(-1j * (w / 2 - k * k / (2 * w)) * t)
        assert abs(c - complex(x.auto_re, x.auto_im)) < 2e-7
        assert abs(x.norm - 1.0) < 1e-8


def test_vmf_matches_cmf2_on_coupled_case(tmp_path):
    """VMF and CMF2 integrate the same equations of motion: on a coupled
    two-state case with one Gaussian mode they agree to the integrators'
    tolerances, and VMF conserves norm and energy."""
    prob = make_problem("g1", str(tmp_path))
    tf = 1.0 / AU2FS
    oc, pc, _ = Oracle(prob).propagate(0.0, tf)
    ov, pv, _ = Oracle(prob).propagate_vmf(2, 0.0, tf)
    for a, b in zip(oc, ov):
        assert abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im)) < 5e-7
        assert abs(b.norm - 1.0) < 1e-7
        assert abs(b.energy - ov[0].energy) < 1e-8
    assert np.abs(pc - pv).max() < 5e-7
