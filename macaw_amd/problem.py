"""Problem descriptor: the module data Macaw holds after ``read_input`` and
the DVR part of ``init_calc`` (source/init_calc.f90:1-226), plus the initial
wavefunction of the ``single-conf`` INIT-WF branch
(source/init_calc.f90:482-752 and ``SetSPF`` :763-927).

All integer tables are int64 and column-major with 1-based entries, exactly
as the Fortran arrays they mirror (``nSPF(mode, state)``, ``iDOF(k, mode)``,
``iOpDof(dof, term)``), so the same buffers can be handed over by the
Fortran driver or by this module.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from .dvr import harm_dvr, param_operator_on_grid, sin_dvr
from .inputfile import InputDeck, InputError, OpParams, parse_input, parse_operator_file


@dataclass
class Problem:
    nstates: int
    nmodes: int
    ndof: int
    nmodeDOF: np.ndarray          # int64[nmodes]
    iDOF: np.ndarray              # int64[maxDOF, nmodes] (Fortran order)
    MultiSet: np.ndarray          # int64[nmodes]
    nSPF: np.ndarray              # int64[nmodes, nstates] (Fortran order)
    ngp: np.ndarray               # int64[ndof]
    grid: np.ndarray              # float64[ngpMax, ndof]
    d1: np.ndarray                # float64[ngpMax, ngpMax, ndof]
    d2: np.ndarray
    xd1: np.ndarray
    ae: np.ndarray                # complex128[nHam]
    lhs_el: np.ndarray            # int64[nHam]
    rhs_el: np.ndarray            # int64[nHam]
    iOpDof: np.ndarray            # int64[ndof, nHam] (Fortran order)
    eps_int: float = 1e-8
    eps_psi: float = 1e-8
    eps_rho: float = 1e-8
    writetime: float = 1.0 / 0.02418884326505
    min_stepsize: float = 1e-9 / 0.02418884326505
    tinit: float = 0.0
    tfinal: float = 0.0
    int_type_psi: int = 1
    int_type: int = 10            # 10 CMF2, 1 VMF Runge-Kutta (Cash-Karp), 2 VMF Dormand-Prince
    normalize: int = 0            # inorm_psi (globals.f90:23-26)
    # parametrised operators (codes 200 + p, -200 - p): the table of their
    # definitions and their values on every DVR DOF's grid
    op_params: list = field(default_factory=list)
    gridoper: np.ndarray = None   # float64[ngpMax, nOpParam, ndof] (Fortran order)
    expect: list = field(default_factory=list)   # [(file, terms)] of `expect = file`
    A0: np.ndarray = None         # complex128[nAConftot]
    psi0: np.ndarray = None       # complex128[dimpsi]
    wgrid: np.ndarray = None
    dof_gwp: np.ndarray = None    # int64[ndof], 1 = Gaussian DOF
    gwp_a: np.ndarray = None      # float64[maxDOF, nSPFmax, nmodes] (Fortran order)
    eps_C: float = 1e-8
    eps_ovl: float = 1e-8
    extra: dict = field(default_factory=dict)

    # ---- derived layout (source/init_calc.f90:25-34, 160-181) ----
    @property
    def nHam(self) -> int:
        return int(self.ae.shape[0])

    @property
    def maxDOF(self) -> int:
        return int(self.iDOF.shape[0])

    @property
    def ngpMax(self) -> int:
        return int(self.grid.shape[0])

    def nAConf(self) -> np.ndarray:
        return np.prod(self.nSPF, axis=0).astype(np.int64)

    def AStateOffset(self) -> np.ndarray:
        n = self.nAConf()
        return np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)

    @property
    def nAConftot(self) -> int:
        return int(self.nAConf().sum())

    def ngpm(self) -> np.ndarray:
        out = np.ones(self.nmodes, dtype=np.int64)
        for m in range(self.nmodes):
            for k in range(int(self.nmodeDOF[m])):
                out[m] *= self.ngp[self.iDOF[k, m] - 1]
        return out

    def gwp_mode(self, m: int) -> bool:
        return self.dof_gwp is not None and bool(self.dof_gwp[self.iDOF[0, m] - 1])

    def npar(self) -> np.ndarray:
        """Parameters per SPF: grid points, or nmodeDOF for a Gaussian mode."""
        ngpm = self.ngpm()
        return np.array([int(self.nmodeDOF[m]) if self.gwp_mode(m) else int(ngpm[m])
                         for m in range(self.nmodes)], dtype=np.int64)

    def ipsiSPF(self):
        """1-based start of each (mode, state) SPF block and dimpsi
        (source/init_calc.f90:160-181)."""
        ip = np.zeros((self.nmodes, self.nstates), dtype=np.int64)
        dim = 0
        ngpm = self.npar()
        for s in range(self.nstates):
            for m in range(self.nmodes):
                if not self.MultiSet[m] and s > 0:
                    ip[m, s] = ip[m, 0]
                    continue
                ip[m, s] = dim + 1
                dim += int(ngpm[m] * self.nSPF[m, s])
        return ip, dim

    @property
    def dimpsi(self) -> int:
        return self.ipsiSPF()[1]


def build_problem(deck: InputDeck, operator_path: str | None = None) -> Problem:
    ns, nm, nd = deck.nstates, deck.nmodes, deck.ndof
    if deck.int_type == 10:
        # the CMF2 step integrates A only by Lanczos and the SPFs only by
        # Cash-Karp / Dormand-Prince; the reference stops on the others when
        # the step runs (cmf2_base.f90:97-100, 252-255)
        if deck.int_type_a != 5:
            raise InputError("Error: integrator not yet implemented for the A coefficients")
        if deck.int_type_psi not in (1, 2):
            raise InputError("Error: integrator not yet implemented for the SPFs")
    if deck.normalize not in (0, 1, 2, 3):
        raise InputError(f"normalize = {deck.normalize}: expected 0, 1, 2 or 3")
    maxdof = max(len(s.dofs) for s in deck.spf)
    iDOF = np.zeros((maxdof, nm), dtype=np.int64, order="F")
    nmodeDOF = np.zeros(nm, dtype=np.int64)
    ms = np.zeros(nm, dtype=np.int64)
    nSPF = np.zeros((nm, ns), dtype=np.int64, order="F")
    for m, s in enumerate(deck.spf):
        nmodeDOF[m] = len(s.dofs)
        iDOF[: len(s.dofs), m] = s.dofs
        ms[m] = 1 if s.multiset else 0
        nSPF[m, :] = s.nspf
    ngp = np.array([deck.grid[k].npts for k in range(1, nd + 1)], dtype=np.int64)
    gmax = int(ngp.max())
    grid = np.zeros((gmax, nd), order="F")
    wgrid = np.zeros((gmax, nd), order="F")
    d1 = np.zeros((gmax, gmax, nd), order="F")
    d2 = np.zeros_like(d1)
    xd1 = np.zeros_like(d1)
    dof_gwp = np.zeros(nd, dtype=np.int64)
    trafos, grids = {}, {}
    for k in range(nd):
        g = deck.grid[k + 1]
        grids[k] = (g.kind, g.npts, g.x0, g.mw)
        if g.kind == "gwp":
            # Gaussian DOF: the declared grid only serves plotting in the reference
            dof_gwp[k] = 1
            grid[:g.npts, k] = np.linspace(g.x0, g.mw, g.npts)
            continue
        if g.kind == "sin":
            x, w, a1, a2, ax, trafos[k] = sin_dvr(g.npts, g.x0, g.mw)
        else:
            x, w, a1, a2, ax, trafos[k] = harm_dvr(g.npts, g.x0, g.mw)
        n = g.npts
        grid[:n, k] = x
        wgrid[:n, k] = w
        d1[:n, :n, k] = a1
        d2[:n, :n, k] = a2
        xd1[:n, :n, k] = ax
    path = operator_path or os.path.join(deck.basedir, deck.hamiltonian)
    registry = OpParams()
    terms = parse_operator_file(_find_file(path), ns, registry)
    nham = len(terms)
    ae = np.array([t.coeff for t in terms], dtype=np.complex128)
    lhs = np.array([t.lhs for t in terms], dtype=np.int64)
    rhs = np.array([t.rhs for t in terms], dtype=np.int64)
    iop = np.zeros((nd, nham), dtype=np.int64, order="F")
    for j, t in enumerate(terms):
        for kd, code in t.ops.items():
            iop[kd - 1, j] = code
    prob = Problem(ns, nm, nd, nmodeDOF, iDOF, ms, nSPF, ngp, grid, d1, d2, xd1,
                   ae, lhs, rhs, iop, eps_int=deck.eps_int, eps_psi=deck.eps_psi,
                   eps_rho=deck.eps_rho, writetime=deck.tout,
                   min_stepsize=deck.min_stepsize, tinit=deck.tinit,
                   tfinal=deck.tfinal, int_type_psi=deck.int_type_psi or 1, int_type=deck.int_type,
                   wgrid=wgrid, eps_C=deck.eps_c, eps_ovl=deck.eps_ovl, normalize=deck.normalize)
    # output files beyond the metric's (macaw_amd/densities.py): the DVR-to-FBR
    # transformations and grid declarations they need
    prob.extra.update(trafo=trafos, grids=grids, hamiltonian_file=deck.hamiltonian,
                      densmat=list(deck.densmat), density=deck.density, wp_ovl=deck.wp_ovl)
    for im in deck.densmat:
        if not 1 <= im <= nm:
            raise InputError(f"densmat = {im}: no such mode")
    for f in deck.expect:
        epath = f if os.path.isabs(f) else os.path.join(deck.basedir, f)
        prob.expect.append((f, parse_operator_file(_find_file(epath), ns, registry)))
    prob.op_params = list(registry.items)
    if prob.op_params:
        # values on every DVR DOF's grid (hamilton.f90:212-300); a
        # parametrised operator on a Gaussian DOF is not supported here
        npar = len(prob.op_params)
        prob.gridoper = np.zeros((gmax, npar, nd), order="F")
        for k in range(nd):
            if dof_gwp[k]:
                continue
            n = int(ngp[k])
            for p_, op in enumerate(prob.op_params):
                prob.gridoper[:n, p_, k] = param_operator_on_grid(op, grid[:n, k], d2[:n, :n, k])
        for _, terms_ in [("", terms)] + prob.expect:
            for t in terms_:
                for kd, c in t.ops.items():
                    if (c > 200 or c <= -200) and dof_gwp[kd - 1]:
                        raise InputError(f"parametrised operator on the Gaussian DOF {kd}: not supported by this engine")
    if dof_gwp.any():
        for m in range(nm):
            kinds = {int(dof_gwp[d - 1]) for d in iDOF[:nmodeDOF[m], m]}
            if len(kinds) > 1:
                raise ValueError("A mode cannot have simultaneously dvr and gwp representation")
        prob.dof_gwp = dof_gwp
        prob.gwp_a = np.zeros((maxdof, int(nSPF.max()), nm), order="F")
    if deck.restart_file:
        from .restart import read_restart, restart_state
        path = deck.restart_file if os.path.isabs(deck.restart_file) else os.path.join(deck.basedir, deck.restart_file)
        prob.A0, prob.psi0 = restart_state(prob, read_restart(path), deck.restart_state, deck.restart_assoc)
    else:
        prob.A0, prob.psi0 = initial_wavefunction(prob, deck)
    return prob


def _find_file(path: str) -> str:
    """The file, or when it does not exist the single file of its directory
    whose name matches it without regard to case (Macaw's own example names
    `Ham_interp.op` for `ham_interp.op`, written on a case-insensitive file
    system); the missing-file error otherwise."""
    if os.path.exists(path):
        return path
    d, b = os.path.split(path)
    try:
        hits = [f for f in os.listdir(d or ".") if f.lower() == b.lower()]
    except OSError:
        hits = []
    return os.path.join(d, hits[0]) if len(hits) == 1 else path


def with_operator(prob: Problem, terms) -> Problem:
    """The problem with its Hamiltonian replaced by the operator `terms` (the
    same layout, state and tables): an expectation operator run through the
    engine's H.A machinery."""
    import dataclasses
    nd = prob.ndof
    ae = np.array([t.coeff for t in terms], dtype=np.complex128)
    lhs = np.array([t.lhs for t in terms], dtype=np.int64)
    rhs = np.array([t.rhs for t in terms], dtype=np.int64)
    iop = np.zeros((nd, len(terms)), dtype=np.int64, order="F")
    for j, t in enumerate(terms):
        for kd, code in t.ops.items():
            iop[kd - 1, j] = code
    return dataclasses.replace(prob, ae=ae, lhs_el=lhs, rhs_el=rhs, iOpDof=iop, expect=[])


def load_problem(input_path: str) -> Problem:
    return build_problem(parse_input(input_path))


def initial_wavefunction(prob: Problem, deck: InputDeck):
    """single-conf initial state: A = e_1 on the chosen state, SPFs from
    ``SetSPF`` (q-multiplication, Gram-Schmidt against SPF 1, then a
    Householder QR of the virtual SPFs), then normalisation
    (source/init_calc.f90:489-497, 525-546, 763-927, 215-221)."""
    ns, nm = prob.nstates, prob.nmodes
    gmax = prob.ngpMax
    phi = np.zeros((gmax, prob.ndof), dtype=np.complex128)
    eps = np.finfo(np.float64).eps
    for k in range(1, prob.ndof + 1):
        spec = deck.init_spf.get(k)
        if spec is None:
            raise ValueError(f"no initial SPF given for dof {k}")
        n = int(prob.ngp[k - 1])
        x = prob.grid[:n, k - 1]
        w = prob.wgrid[:n, k - 1]
        if abs(spec.x0) < eps and abs(spec.p0) < eps:
            for j in range(1, n // 2 + 2):
                phi[j - 1, k - 1] = np.exp(-(0.5 * x[j - 1] / spec.dx) ** 2) * np.sqrt(w[j - 1])
                phi[n - j, k - 1] = phi[j - 1, k - 1]
        else:
            xx = x - spec.x0
            phi[:n, k - 1] = np.exp(-(0.5 * xx / spec.dx) ** 2 + 1j * spec.p0 * xx) * np.sqrt(w)
    ip, dimpsi = prob.ipsiSPF()
    psi = np.zeros(dimpsi, dtype=np.complex128)
    ngpm = prob.ngpm()
    for m in range(nm):
        if prob.gwp_mode(m):
            gaussian_lattice(prob, deck, m, psi, ip)
            continue
        nmd = int(prob.nmodeDOF[m])
        dofs = [int(d) for d in prob.iDOF[:nmd, m]]
        sizes = [int(prob.ngp[d - 1]) for d in dofs]
        ngpt = int(ngpm[m])
        nS = int(prob.nSPF[m, :].max())
        # product grid, first DOF fastest
        idx = np.indices(sizes[::-1]).reshape(nmd, -1)[::-1]
        psi_t = np.zeros((ngpt, nS), dtype=np.complex128)
        col = np.ones(ngpt, dtype=np.complex128)
        for k in range(nmd):
            col = col * phi[idx[k], dofs[k] - 1]
        psi_t[:, 0] = col / np.linalg.norm(col)
        kg = 0
        jj = 1
        for j in range(2, nS + 1):
            jj += 1
            kg = kg % nmd + 1
            lref = (jj - 2) // nmd + 1
            op = prob.grid[idx[kg - 1], dofs[kg - 1] - 1]
            psi_t[:, j - 1] = psi_t[:, lref - 1] * op
        for j in range(1, nS):
            c = np.vdot(psi_t[:, 0], psi_t[:, j])
            psi_t[:, j] = psi_t[:, j] - c * psi_t[:, 0]
        if nS > 1:
            q, _ = np.linalg.qr(psi_t[:, 1:], mode="reduced")
            psi_t[:, 1:] = q
        for s in range(ns):
            n_s = int(prob.nSPF[m, s])
            start = ip[m, s] - 1
            psi[start:start + ngpt * n_s] = psi_t[:, :n_s].reshape(-1, order="F")
    A = np.zeros(prob.nAConftot, dtype=np.complex128)
    A[prob.AStateOffset()[deck.init_state - 1]] = 1.0
    # norm_wf of a single populated configuration: product of the first-SPF norms
    cnorm = 1.0 + 0.0j
    s = deck.init_state - 1
    for m in range(nm):
        if prob.gwp_mode(m):
            continue            # normalised Gaussians
        n = int(ngpm[m])
        v = psi[ip[m, s] - 1: ip[m, s] - 1 + n]
        cnorm = cnorm * np.vdot(v, v)
    A = A / np.sqrt(cnorm.real)
    return A, psi


def gaussian_lattice(prob: Problem, deck: InputDeck, m: int, psi: np.ndarray, ip) -> None:
    """Initial Gaussians of mode m (``SetSPF``, source/init_calc.f90:958-1081):
    widths w_l = dx scw^|l| and centres xi_l along each DOF from the overlap
    between neighbours, lattice indices by neighbour shells (distribution 0)
    or from the file gwpgrid_<mode> (distribution 1); A = -(1/(2 w))^2 and
    xi written for every electronic state."""
    nmd = int(prob.nmodeDOF[m])
    dofs = [int(d) for d in prob.iDOF[:nmd, m]]
    nmax = int(prob.nSPF.max())
    xi_v = np.zeros((2 * nmax + 1, nmd), dtype=np.complex128)
    w_v = np.zeros((2 * nmax + 1, nmd))
    dists = set()
    for j, k in enumerate(dofs):
        spec = deck.init_spf.get(k)
        if spec is None or spec.kind != "gwp":
            raise ValueError(f"Gaussian DOF {k} needs a gwp initial SPF")
        dists.add(spec.dist)
        dx, scw = spec.dx, spec.scw
        for l in range(-nmax, nmax + 1):
            w_v[l + nmax, j] = dx * scw ** abs(l)
        xi_v[nmax, j] = 0.0
        for l in range(1, nmax + 1):
            d = -np.log(spec.ovl) + 0.5 * np.log(2 * scw / (1.0 + scw ** 2))
            d = d * (1.0 + scw ** 2) / w_v[l - 1 + nmax, j] ** 2
            d = (xi_v[l - 1 + nmax, j].real + np.sqrt(d)) / scw ** 2
            xi_v[l + nmax, j] = d
            xi_v[-l + nmax, j] = -d
        eps = np.finfo(np.float64).eps
        if max(abs(spec.x0), abs(spec.p0)) < eps:
            phi = 0.0
        else:
            phi = complex(0.5 * spec.x0 / spec.dx ** 2, spec.p0)
        if abs(phi) > eps:
            xi_v[:, j] += phi
    if len(dists) > 1:
        raise ValueError(f"the same GWP distribution must be used for all DOFs of mode {m + 1}")
    dist = dists.pop()
    nS = int(prob.nSPF[m, :].max())
    lat = np.zeros((nmd, nS), dtype=np.int64)
    if nS > 1:
        if dist == 0:
            l = 1
            i1, i2 = 0, 0
            while l < nS:
                for i3 in range(i1, i2 + 1):
                    for j in range(nmd):
                        for step in (-1, 1):
                            cand = lat[:, i3].copy()
                            cand[j] += step
                            if any(np.array_equal(cand, lat[:, q]) for q in range(l)):
                                continue
                            lat[:, l] = cand
                            l += 1
                            if l >= nS:
                                break
                        if l >= nS:
                            break
                    if l >= nS:
                        break
                i1, i2 = i2 + 1, l - 1
        elif dist == 1:
            path = os.path.join(deck.basedir, f"gwpgrid_{m + 1}")
            rows = np.loadtxt(path, dtype=np.int64, ndmin=2)
            lat[:, :] = rows[:nS, :nmd].T
        else:
            raise ValueError(f"Unknown distribution type: {dist}")
    xi = np.zeros((nmd, nS), dtype=np.complex128)
    for l in range(nS):
        for j in range(nmd):
            xi[j, l] = xi_v[lat[j, l] + nmax, j]
            prob.gwp_a[j, l, m] = -(0.5 / w_v[lat[j, l] + nmax, j]) ** 2
    for s in range(prob.nstates):
        n_s = int(prob.nSPF[m, s])
        start = ip[m, s] - 1
        psi[start:start + nmd * n_s] = xi[:, :n_s].reshape(-1, order="F")

