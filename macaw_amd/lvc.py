"""Seeded linear-vibronic-coupling (LVC) model generator.

Writes Macaw-format input and operator files for the benchmark workloads:

* ``c1``: 2 states, 5 modes, HO-DVR 24 points, 6 SPFs, 100 fs.
* ``c2``: 4 states, 8 modes (+ diagonal quadratic terms), HO-DVR 32 points,
  8 SPFs (8^8 configurations per state), 200 fs.
* ``c3``: G-MCTDH, 3 states, 6 modes; modes 1-3 with 5 Gaussian SPFs (HO
  ground-state widths on the reference's neighbour-shell lattice), modes 4-6
  HO-DVR 28 points, 7 SPFs; C-matrix Tikhonov eps 1e-8, CMF2 eps 1e-5 (the
  DeltaPsi estimate of these wavefunctions has a rounding floor near 1e-7),
  150 fs.
* ``t*``, ``g*``: small variants used by the tests (``g*`` with Gaussian modes).

Energies are drawn in eV with ``numpy.random.default_rng(seed)`` and written
in Hartree (divided by Macaw's ``au2eV = 27.2113961``).  Hamiltonian lines:
``-w/2 | 1 |i dq^2``, ``w/2 | 1 |i q^2``, ``E_s | Ss&s``, ``kappa | Ss&s |i q``,
``gamma | Ss&s |i q^2`` and ``lambda | Sa&b |j q``.
"""
from __future__ import annotations

import itertools
import os
from dataclasses import dataclass

import numpy as np

from .inputfile import AU2EV


@dataclass
class LVCSpec:
    name: str
    seed: int
    nstates: int
    nmodes: int
    npts: int
    nspf: object                 # int, or one count per electronic state
    tfinal: float
    tout: float = 0.5
    omega: tuple = (0.02, 0.40)
    kappa_sd: float = 0.10
    gamma_sd: float = 0.0
    n_lambda: int = 1
    lam: tuple = (0.10, 0.10)
    gaps: tuple = ()
    eps_int: float = 1e-7
    eps_psi: float = 1e-7
    spf_integrator: str = "rk45"
    init_state: int = 0          # 0 -> top state
    multiset: bool = True
    # G-MCTDH: the first gwp_modes modes carry ngauss Gaussian SPFs
    gwp_modes: int = 0
    ngauss: int = 5
    gwp_ovl: float = 0.7
    gwp_scw: float = 1.0
    gwp_multiset: bool = True
    eps_C: float = 1e-8
    eps_S: float = 1e-8


SPECS = {
    "c1": LVCSpec("c1", 101, 2, 5, 24, 6, 100.0, omega=(0.02, 0.40), kappa_sd=0.10,
                  n_lambda=1, lam=(0.10, 0.10), gaps=(-0.25, 0.25)),
    "c2": LVCSpec("c2", 202, 4, 8, 32, 8, 200.0, omega=(0.05, 0.40), kappa_sd=0.12,
                  gamma_sd=0.02, n_lambda=6, lam=(0.05, 0.20),
                  gaps=(0.0, 0.3, 0.6, 0.9)),
    # small cases for the tests
    "t1": LVCSpec("t1", 11, 2, 3, 8, 3, 5.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  n_lambda=1, lam=(0.10, 0.10), gaps=(-0.2, 0.2)),
    "t2": LVCSpec("t2", 12, 3, 3, 7, (3, 4, 2), 5.0, omega=(0.05, 0.30), kappa_sd=0.12,
                  gamma_sd=0.02, n_lambda=3, lam=(0.05, 0.20), gaps=(0.0, 0.25, 0.5)),
    "t3": LVCSpec("t3", 13, 2, 4, 10, 4, 10.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  gamma_sd=0.02, n_lambda=2, lam=(0.05, 0.15), gaps=(-0.1, 0.1),
                  spf_integrator="dopri"),
    "t4": LVCSpec("t4", 14, 2, 3, 8, 3, 5.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  n_lambda=1, lam=(0.10, 0.10), gaps=(-0.2, 0.2), multiset=False),
    # 8 SPFs everywhere (the 8x8 matrix-core path), two mode groups
    "t5": LVCSpec("t5", 15, 3, 4, 12, 8, 5.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  gamma_sd=0.02, n_lambda=3, lam=(0.05, 0.15), gaps=(-0.2, 0.0, 0.2)),
    # G-MCTDH: 3 states, modes 1-3 with 5 Gaussians, modes 4-6 HO-DVR 28 / 7 SPFs
    "c3": LVCSpec("c3", 303, 3, 6, 28, 7, 150.0, omega=(0.05, 0.35), kappa_sd=0.10,
                  n_lambda=3, lam=(0.05, 0.15), gaps=(-0.2, 0.0, 0.2), gwp_modes=3, ngauss=5,
                  eps_C=1e-8, eps_int=1e-5),
    # small G-MCTDH cases for the tests
    "g1": LVCSpec("g1", 21, 2, 3, 8, 3, 5.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  n_lambda=1, lam=(0.10, 0.10), gaps=(-0.2, 0.2), gwp_modes=1, ngauss=3, eps_int=1e-6),
    "g2": LVCSpec("g2", 22, 2, 3, 8, 3, 5.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  gamma_sd=0.02, n_lambda=2, lam=(0.05, 0.15), gaps=(-0.1, 0.1), gwp_modes=2,
                  ngauss=3, gwp_multiset=False, spf_integrator="dopri", eps_int=1e-6),
    "g3": LVCSpec("g3", 23, 3, 4, 8, 3, 5.0, omega=(0.05, 0.30), kappa_sd=0.10,
                  n_lambda=3, lam=(0.05, 0.15), gaps=(-0.2, 0.0, 0.2), gwp_modes=2, ngauss=4, eps_int=1e-6),
}


def draw(spec: LVCSpec):
    """Seeded parameters in eV: omega[f], kappa[S,f], gamma[S,f], couplings."""
    rng = np.random.default_rng(spec.seed)
    omega = rng.uniform(spec.omega[0], spec.omega[1], size=spec.nmodes)
    kappa = rng.normal(0.0, spec.kappa_sd, size=(spec.nstates, spec.nmodes))
    gamma = (rng.normal(0.0, spec.gamma_sd, size=(spec.nstates, spec.nmodes))
             if spec.gamma_sd > 0 else np.zeros((spec.nstates, spec.nmodes)))
    pairs = list(itertools.combinations(range(1, spec.nstates + 1), 2))
    combos = [(m, p) for m in range(1, spec.nmodes + 1) for p in pairs]
    couplings = []
    if spec.n_lambda == 1 and spec.lam[0] == spec.lam[1]:
        j = int(rng.integers(1, spec.nmodes + 1))
        couplings.append((j, pairs[0], spec.lam[0]))
    else:
        pick = rng.choice(len(combos), size=min(spec.n_lambda, len(combos)), replace=False)
        for c in pick:
            j, p = combos[int(c)]
            couplings.append((j, p, float(rng.uniform(spec.lam[0], spec.lam[1]))))
    return omega, kappa, gamma, couplings


def _ev(x: float) -> str:
    return f"{x / AU2EV: .12e}".replace("e", "d")


def write_lvc(spec: LVCSpec, outdir: str) -> str:
    """Write ``<name>.inp`` and ``<name>.op``; return the input path."""
    os.makedirs(outdir, exist_ok=True)
    omega, kappa, gamma, couplings = draw(spec)
    S, f = spec.nstates, spec.nmodes
    op = [f"# LVC model {spec.name}, seed {spec.seed}; energies in Hartree"]
    for i in range(f):
        op.append(f"{_ev(-omega[i] / 2)}  |  1  |{i + 1:<2d} dq^2")
        op.append(f"{_ev(omega[i] / 2)}  |  1  |{i + 1:<2d} q^2")
    for s in range(S):
        gap = spec.gaps[s] if spec.gaps else 0.0
        op.append(f"{_ev(gap)}  |  S{s + 1}&{s + 1}")
    for s in range(S):
        for i in range(f):
            op.append(f"{_ev(kappa[s, i])}  |  S{s + 1}&{s + 1}  |{i + 1:<2d} q")
    if spec.gamma_sd > 0:
        for s in range(S):
            for i in range(f):
                op.append(f"{_ev(gamma[s, i])}  |  S{s + 1}&{s + 1}  |{i + 1:<2d} q^2")
    for j, (a, b), lam in couplings:
        op.append(f"{_ev(lam)}  |  S{a}&{b}  |{j:<2d} q")
    op.append("----")
    opname = f"{spec.name}.op"
    with open(os.path.join(outdir, opname), "w") as fh:
        fh.write("\n".join(op) + "\n")
    top = spec.init_state or S
    lines = [f"tfinal = {spec.tfinal}", f"tout = {spec.tout}", "",
             f"nstates = {S}", f"ndof = {f}", f"nmodes = {f}", "", "steps", "", "SPF"]
    for i in range(f):
        gauss = i < spec.gwp_modes
        multi = spec.gwp_multiset if gauss else spec.multiset
        if gauss:
            counts = [spec.ngauss] * S
        else:
            counts = list(spec.nspf) if isinstance(spec.nspf, (list, tuple)) else [spec.nspf] * S
        lines.append(f" {'M' if multi else 'S'}, {i + 1} = " + ",".join(str(c) for c in counts))
    lines += ["", "GRID"]
    for i in range(f):
        if i < spec.gwp_modes:
            lines.append(f"  {i + 1}  gwp  41  -8.0  8.0")
        else:
            lines.append(f"  {i + 1}  HO  {spec.npts}  0.0  1.0")
    lines += ["", f"Hamiltonian = {opname}", "", "INIT-WF", "   single-conf", f"   el  {top}"]
    for i in range(f):
        if i < spec.gwp_modes:
            lines.append(f"   {i + 1}   gwp 0  0.0  0.0  0.707107    ovl = {spec.gwp_ovl}   scw = {spec.gwp_scw}")
        else:
            lines.append(f"   {i + 1}   gauss  0.0  0.0  0.707107")
    fd = lambda x: f"{x:.1e}".replace("e", "d")
    lines += ["END-INIT-WF", "", f"eps_int = {fd(spec.eps_int)}"]
    if spec.gwp_modes:
        lines += [f"eps_C = {fd(spec.eps_C)}", f"eps_S = {fd(spec.eps_S)}"]
    lines += ["", "integration = CMF2", "Lanczos  1.d-7",
              f"{spec.spf_integrator}     {fd(spec.eps_psi)}", ""]
    path = os.path.join(outdir, f"{spec.name}.inp")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def make_problem(spec: LVCSpec | str, outdir: str):
    from .problem import load_problem
    if spec == "c3j":
        return make_c3_jittered(outdir)
    if isinstance(spec, str):
        spec = SPECS[spec]
    return load_problem(write_lvc(spec, outdir))


def make_c3_jittered(outdir: str, eps_int: float = 1e-7, sd: float = 0.2):
    """C3 as the workload states it: the Gaussian SPFs of modes 1-3 keep the
    HO ground-state width (A = -1/2) and get centres q0 ~ N(0, sd) (for
    A = -1/2, xi = -2 A q0 = q0), drawn from the generator seeded with
    (303, 1) so the Hamiltonian draws of seed 303 are unchanged; one draw
    per (state, mode, Gaussian).  The initial state (HO ground state on the
    top state) is the least-squares projection of the ground-state Gaussian
    on each mode's jittered Gaussians, A = product of the per-mode
    coefficients on the top state, normalised.  Macaw's input has no way to
    place Gaussians at random, so the state goes through a Macaw `restart`
    file (output.f90:230-241) read by an INIT-WF `restart` block
    (init_calc.f90:290-481), the route the reference itself offers."""
    import dataclasses
    from .problem import load_problem
    from .restart import gaussian_overlap, trace_norm, write_restart
    spec = dataclasses.replace(SPECS["c3"], name="c3j", eps_int=eps_int)
    base_path = write_lvc(spec, outdir)
    prob = load_problem(base_path)
    rng = np.random.default_rng([spec.seed, 1])
    ip, _ = prob.ipsiSPF()
    npar = prob.npar()
    psi = prob.psi0.copy()
    top = spec.nstates - 1 if spec.init_state == 0 else spec.init_state - 1
    off = prob.AStateOffset()
    coef = []
    for m in range(prob.nmodes):
        if not prob.gwp_mode(m):
            continue
        for st in range(prob.nstates):
            if st > 0 and not prob.MultiSet[m]:
                continue
            n = int(prob.nSPF[m, st])
            nmd = int(npar[m])
            start = int(ip[m, st]) - 1
            xi = (rng.normal(0.0, sd, size=(nmd, n)) + 0j) * (-2.0 * prob.gwp_a[:nmd, :n, m])
            psi[start:start + nmd * n] = xi.reshape(-1, order="F")
            if st == top or (st == 0 and not prob.MultiSet[m]):
                # projection of the ground-state Gaussian (xi = 0) on the set
                xa = np.concatenate([xi, np.zeros((nmd, 1), complex)], axis=1)
                S6 = _overlap_with_ground(prob, m, xa)
                c = np.linalg.solve(S6[:n, :n], S6[:n, n])
                coef.append((m, c))
    A = np.zeros(prob.nAConftot, np.complex128)
    dims = [int(prob.nSPF[m, top]) for m in range(prob.nmodes)]
    a = np.ones(1, complex)
    for m in range(prob.nmodes):
        v = np.zeros(dims[m], complex)
        cm = [c for mm, c in coef if mm == m]
        if cm:
            v[:] = cm[0]
        else:
            v[0] = 1.0     # grid modes: the ground state is the first SPF
        a = np.kron(v, a)  # mode 1 fastest
    A[off[top]: off[top] + a.size] = a
    A /= np.sqrt(trace_norm(prob, A, psi))
    rpath = os.path.join(outdir, "c3j.restart")
    write_restart(rpath, prob, 0.0, A, psi)
    text = open(base_path).read()
    i, j = text.index("INIT-WF"), text.index("END-INIT-WF")
    assoc = "\n".join(f"   {s + 1} <-- {s + 1}" for s in range(prob.nstates))
    text = text[:i] + f"INIT-WF\n   restart\n   {os.path.basename(rpath)}\n   0\n{assoc}\n" + text[j:]
    path = os.path.join(outdir, "c3j.inp")
    with open(path, "w") as fh:
        fh.write(text)
    return load_problem(path)


def _overlap_with_ground(prob, m, xi):
    """Normalised overlaps of the Gaussians xi[:, :n] of mode m and one more
    Gaussian (the last column) of the same width (the mode's first width)."""
    from .restart import gaussian_overlap
    nmd, n1 = xi.shape
    a = prob.gwp_a
    saved = a[:nmd, :n1, m].copy() if a.shape[1] >= n1 else None
    if a.shape[1] < n1:
        grown = np.zeros((a.shape[0], n1, a.shape[2]), order="F")
        grown[:, :a.shape[1], :] = a
        grown[:nmd, a.shape[1]:, m] = a[:nmd, :1, m]
        tmp = prob.gwp_a
        prob.gwp_a = grown
        try:
            return gaussian_overlap(prob, m, xi)
        finally:
            prob.gwp_a = tmp
    prob.gwp_a[:nmd, n1 - 1, m] = prob.gwp_a[:nmd, 0, m]
    try:
        return gaussian_overlap(prob, m, xi)
    finally:
        prob.gwp_a[:nmd, :n1, m] = saved
