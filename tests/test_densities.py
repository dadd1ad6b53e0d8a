"""Macaw's densmat / density / wpOvl / log files (macaw_amd/densities.py).

Pins (CPU): for a product SPF f(x1) g(x2) of a two-DOF HO mode the reduced
density along each DOF is |f|^2 and |g|^2 (a transposed grid index fails
this); the HO ground state's basis populations are (1, 0, ...); the sums of
Density and DensityPop are tr(rho S) for non-orthonormal SPFs; only the
upper triangle of Macaw's rho is read; a Gaussian DOF's density is the
closed-form |g(x)|^2 times the grid step; the record layouts read back; the
text lines equal the reference's edit descriptors printed by flang.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from macaw_amd import densities as D
from macaw_amd.problem import load_problem
from macaw_amd.restart import _read_record

FLANG = "/opt/rocm/lib/llvm/bin/flang"

DECK = """tfinal = 1.0
tout = 0.5

nstates = 2
ndof = 3
nmodes = 2

densmat = 2
densmat = 1
density
wp_ovl

SPF
 M, 1, 2 = 2,2
 M, 3 = 2,2

GRID
  1  HO  6  0.0  1.0
  2  HO  5  0.3  0.8
  3  gwp  41  -6.0  6.0

Hamiltonian = h.op

INIT-WF
   single-conf
   el  1
   1   gauss  0.0  0.0  0.707107
   2   gauss  0.0  0.0  0.707107
   3   gwp 0  0.0  0.0  0.707107
END-INIT-WF

integration = CMF2
Lanczos  1.d-7
rk45     1.0d-07
"""

HAM = """ 1.0d-03  |  1  |1  q^2
 1.0d-03  |  1  |2  q^2
 1.0d-03  |  1  |3  q^2
-----
"""


@pytest.fixture
def prob(tmp_path):
    (tmp_path / "p.inp").write_text(DECK)
    (tmp_path / "h.op").write_text(HAM)
    return load_problem(str(tmp_path / "p.inp"))


def _rho_engine(R):
    """The engine's rho[s, m] layout (the transpose of Macaw's rho(j1, j2))."""
    return R.T


def _set_block(prob, psi, m, s, blk):
    ip, _ = prob.ipsiSPF()
    jel = s if prob.MultiSet[m] else 0
    start = int(ip[m, jel]) - 1
    psi[start:start + blk.size] = blk


def test_keywords_are_parsed(prob):
    ex = prob.extra
    assert ex["densmat"] == [2, 1] and ex["density"] and ex["wp_ovl"]
    assert ex["grids"][1] == ("HO", 5, 0.3, 0.8) and ex["grids"][2][0] == "gwp"
    assert set(ex["trafo"]) == {0, 1}


def test_product_spf_densities_per_dof(prob):
    """SPF 1 of mode 1 = f(x1) g(x2) (DOF 1 fastest in the block); rho = 1 on
    it: Density along DOF 1 is |f|^2, along DOF 2 |g|^2."""
    rng = np.random.default_rng(3)
    n1, n2 = 6, 5
    f = rng.normal(size=n1) + 1j * rng.normal(size=n1)
    g = rng.normal(size=n2) + 1j * rng.normal(size=n2)
    f /= np.linalg.norm(f)
    g /= np.linalg.norm(g)
    psi = np.zeros(prob.dimpsi, complex)
    for s in range(2):
        spf1 = np.outer(g, f).reshape(-1)            # index i1 + n1 * i2
        spf2 = np.zeros(n1 * n2, complex)
        spf2[3] = 1.0
        _set_block(prob, psi, 0, s, np.concatenate([spf1, spf2]))
    rho = np.zeros((2, 2, 2, 2), complex)
    for s in range(2):
        rho[s, 0] = _rho_engine(np.diag([1.0, 0.0]).astype(complex))
    T = prob.extra["trafo"]
    d1, p1 = D.calc_density(prob, rho, psi, 0, T[0])
    d2, p2 = D.calc_density(prob, rho, psi, 1, T[1])
    for s in range(2):
        assert np.allclose(d1[:, s], np.abs(f) ** 2, atol=1e-14)
        assert np.allclose(d2[:, s], np.abs(g) ** 2, atol=1e-14)
        # basis populations: |T f|^2 (T orthogonal: they add up to 1)
        assert np.allclose(p1[:, s], np.abs(T[0] @ f) ** 2, atol=1e-14)
        assert abs(p2[:, s].sum() - 1.0) < 1e-13


def test_ho_ground_state_basis_populations(prob):
    """The HO ground state in DVR amplitudes is row 0 of the DVR-to-FBR
    transformation: its basis populations are (1, 0, ..., 0)."""
    T = prob.extra["trafo"]
    psi = np.zeros(prob.dimpsi, complex)
    for s in range(2):
        _set_block(prob, psi, 0, s, np.concatenate([np.outer(T[1][0], T[0][0]).reshape(-1),
                                                    np.zeros(30, complex)]))
    rho = np.zeros((2, 2, 2, 2), complex)
    rho[:, 0] = _rho_engine(np.diag([1.0, 0.0]).astype(complex))
    dens, pops = D.calc_density(prob, rho, psi, 0, T[0])
    assert np.allclose(pops[:, 0], np.eye(6)[0], atol=1e-13)
    assert np.allclose(dens[:, 0], T[0][0] ** 2, atol=1e-14)


def test_trace_identity_and_upper_triangle(prob):
    """sum_g Density = sum_g DensityPop = tr(rho S) with S the SPF overlaps
    (non-orthonormal SPFs, Hermitian rho); a lower triangle of rubbish in
    Macaw's rho(j1, j2) changes nothing."""
    rng = np.random.default_rng(5)
    psi = rng.normal(size=prob.dimpsi) + 1j * rng.normal(size=prob.dimpsi)
    T = prob.extra["trafo"]
    rho = np.zeros((2, 2, 2, 2), complex)
    Rs = []
    for s in range(2):
        X = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        R = X @ X.conj().T
        Rs.append(R)
        junk = R.copy()
        junk[1, 0] = 17.0 - 4.0j                       # below the diagonal: never read
        rho[s, 0] = _rho_engine(junk)
    for dof in (0, 1):
        dens, pops = D.calc_density(prob, rho, psi, dof, T[dof])
        for s in range(2):
            _, npts, blk = D._mode_block(prob, psi, 0, s)
            Phi = blk.reshape(2, npts)
            S = Phi.conj() @ Phi.T                     # S[j1, j2] = <phi_j1|phi_j2>
            tr = np.einsum("ab,ab->", Rs[s], S).real
            assert abs(dens[:, s].sum() - tr) < 1e-10 * abs(tr)
            assert abs(pops[:, s].sum() - tr) < 1e-10 * abs(tr)


def test_gaussian_density_closed_form(prob):
    """One populated Gaussian exp(a x^2 + xi x): density |N|^2 exp(2 a x^2 +
    2 Re(xi) x) dx on the declared grid, integrating to 1."""
    a, xi = -0.7, 0.4 + 1.3j
    prob.gwp_a[0, 0, 1] = a
    prob.gwp_a[0, 1, 1] = -0.5
    psi = np.zeros(prob.dimpsi, complex)
    for s in range(2):
        _set_block(prob, psi, 1, s, np.array([xi, -1.0 + 0.2j]))
    rho = np.zeros((2, 2, 2, 2), complex)
    rho[:, 1] = _rho_engine(np.diag([1.0, 0.0]).astype(complex))
    dens, pops = D.calc_density(prob, rho, psi, 2)
    assert pops is None
    x = np.linspace(-6.0, 6.0, 41)
    dx = x[1] - x[0]
    # |N|^2 = sqrt(-2a/pi) exp(Re(xi)^2 / (2a))
    ref = np.sqrt(-2 * a / np.pi) * np.exp(xi.real ** 2 / (2 * a)) * np.exp(2 * a * x ** 2 + 2 * xi.real * x) * dx
    assert np.allclose(dens[:, 0], ref, rtol=1e-12, atol=1e-16)
    assert abs(dens[:, 0].sum() - 1.0) < 1e-6


def test_densmat_and_density_records(prob, tmp_path):
    rng = np.random.default_rng(9)
    psi = rng.normal(size=prob.dimpsi) + 1j * rng.normal(size=prob.dimpsi)
    rho = rng.normal(size=(2, 2, 2, 2)) + 1j * rng.normal(size=(2, 2, 2, 2))
    prob.gwp_a[:1, :2, 1] = -0.5
    path = tmp_path / "densmat"
    with open(path, "wb") as fh:
        D.write_densmat(fh, prob, 12.5, rho, psi, [2, 1])
    with open(path, "rb") as fh:
        assert struct.unpack("<d", _read_record(fh)) == (12.5,)
        for im in (2, 1):
            for s in (1, 2):
                rec = _read_record(fh)
                m, st = struct.unpack("<qq", rec[:16])
                assert (m, st) == (im, s)
                r = np.frombuffer(rec[16:16 + 64], "<c16").reshape(2, 2, order="F")
                assert np.array_equal(r, rho[s - 1, im - 1].T)
                _, npts, blk = D._mode_block(prob, psi, im - 1, s - 1)
                assert np.array_equal(np.frombuffer(rec[80:], "<c16"), blk)
                assert len(rec) == 80 + 16 * 2 * npts
    recs, lines = D.density_records(prob, 1.0, rho, psi, prob.extra["trafo"])
    # time, 2 DVR DOFs x 2 states (density + basis populations), 1 Gaussian DOF x 2 states
    assert len(recs) == 1 + 4 + 2
    assert len(recs[1]) == 8 + 2 * 8 * 6 and len(recs[3]) == 8 + 2 * 8 * 5 and len(recs[5]) == 8 + 8 * 41
    assert lines[0] == " Population of the basis set underlying the DVR" and lines[1] == "   DOF: 1"


FORTRAN = """
program fmt
  implicit none
  integer :: iel, i, nG
  double precision :: dp(6), t
  complex(8) :: swp
  character(6) :: dvrt
  character(len = 100) :: hf
  read(*,*) dp, t, swp
  write(*,'(a,i2)') '   DOF:', 3
  iel = 2
  nG = 6
  write(*,'(a,i2,a,3x,f7.4,2x,f7.4,a,f7.4,2x,f7.4)') '     State ', iel, ':', dp(1:2), ' ... ', dp(nG - 1:nG)
  write(*,'(a,i2,a,3x,f7.4,2x,f7.4)') '     State ', iel, ':', dp(1), dp(nG)
  write(*,'(a,f7.2,a)') 'Time: ', t, ' fs'
  write(*,'(7x,6(8x,i3,8x))') (i, i = 1, 1)
  write(*,'(i3,2x,6(2x,"(",f7.4,",",f7.4,")"))') 1, swp
  write(*,'(a,i0)') 'Number of electronic states: ', 2
  write(*,*)
  write(*,*) 'SPF structure'
  write(*,'(3x,a,i0,a,2(i0,2x),a,2(i0,2x))') 'Mode ', 1, ' -     DOFs: ', 1, 2, ' ;    SPFs: ', 2, 2
  dvrt = 'HO'
  write(*,'(3x,a,i0,a,a,3x,i0,2(4x,f10.4))') 'DOF ', 2, ' - ', dvrt, 5, 0.3d0, 0.8d0
  hf = 'h.op'
  write(*,*) 'The Hamiltonian is read from file: ', hf
end program fmt
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_lines_match_fortran_edit_descriptors(prob, tmp_path):
    src = tmp_path / "fmt.f90"
    src.write_text(FORTRAN)
    exe = tmp_path / "fmt"
    subprocess.check_call([FLANG, "-O0", "-o", str(exe), str(src)], cwd=str(tmp_path))
    dp = [0.8123456, 0.1, 0.05, 0.02, 0.0125, 0.00371]
    t, swp = 12.25, complex(0.99999991, 0.0)
    text = " ".join(repr(x) for x in dp) + f" {t!r} ({swp.real!r},{swp.imag!r})\n"
    out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    pops = np.array(dp)[:, None] * np.ones((1, 2))
    mine = [f"   DOF:{3:2d}",
            f"     State {2:2d}:   {pops[0, 1]:7.4f}  {pops[1, 1]:7.4f} ... {pops[-2, 1]:7.4f}  {pops[-1, 1]:7.4f}",
            f"     State {2:2d}:   {pops[0, 1]:7.4f}  {pops[-1, 1]:7.4f}"]
    # the same lines as density_records writes them
    recs, lines = D.density_records(prob, 0.0, np.zeros((2, 2, 2, 2), complex), np.zeros(prob.dimpsi, complex),
                                    prob.extra["trafo"])
    assert lines[1] == "   DOF: 1" and lines[2].startswith("     State  1:   ")
    assert out[:3] == mine
    assert out[3:6] == D.wpovl_lines(t / D.AU2FS, swp)
    log = D.log_tables(prob, prob.extra["grids"], "h.op")
    assert out[6] == log[0]
    assert out[7:9] == log[3:5]
    assert out[9] == log[5]
    assert out[10] == log[10]
    assert out[11].rstrip() == log[-1]


@pytest.mark.gpu
def test_run_writes_density_files(tmp_path):
    """The standalone run on t1 with densmat / density / wp_ovl: one densmat
    and density block per output time; every DOF's density adds up to the
    state population the run reports, as do the basis populations; wpOvl
    holds the norm."""
    from macaw_amd import run as RUN
    from macaw_amd.lvc import make_problem
    d = str(tmp_path)
    make_problem("t1", d)
    inp = open(os.path.join(d, "t1.inp")).read().replace("tfinal = 5.0", "tfinal = 1.0")
    inp = inp.replace("steps\n", "densmat = 2\ndensity\nwp_ovl\n")
    with open(os.path.join(d, "d.inp"), "w") as fh:
        fh.write(inp)
    out = os.path.join(d, "out")
    outs, pops = RUN.run(os.path.join(d, "d.inp"), out, log=lambda m: None)
    n = len(outs)
    assert n == 3
    with open(os.path.join(out, "densmat"), "rb") as fh:
        for k in range(n):
            (t,) = struct.unpack("<d", _read_record(fh))
            assert abs(t - outs[k].time) < 1e-9
            for s in (1, 2):
                rec = _read_record(fh)
                assert struct.unpack("<qq", rec[:16]) == (2, s)
                r = np.frombuffer(rec[16:16 + 16 * 9], "<c16").reshape(3, 3, order="F")
                assert abs(np.trace(r).real - pops[k][s - 1]) < 1e-9
    with open(os.path.join(out, "density"), "rb") as fh:
        for k in range(n):
            _read_record(fh)
            for dof in (1, 2, 3):
                for s in (1, 2):
                    rec = _read_record(fh)
                    assert struct.unpack("<q", rec[:8]) == (dof,)
                    v = np.frombuffer(rec[8:], "<f8")
                    dens, bpop = v[:8], v[8:]
                    assert abs(dens.sum() - pops[k][s - 1]) < 1e-9
                    assert abs(bpop.sum() - pops[k][s - 1]) < 1e-9
    text = open(os.path.join(out, "output")).read()
    assert text.count("Population of the basis set underlying the DVR") == n
    w = open(os.path.join(out, "wpOvl")).read().splitlines()
    assert w[0] == "Wave packet overlap matrix" and len(w) == 1 + 3 * n
    assert w[3].startswith("  1    ( 1.0000, 0.0000)")
    assert open(os.path.join(out, "log")).read().startswith("Number of electronic states: 2\n")
