"""Macaw's remaining output files of a propagation, computed on the host
from the state the engine returns (A, psi) and its density matrices:

* ``densmat`` (``densmat = m`` in the input, once per mode): per output time
  an unformatted record with the time, then per listed mode and state one
  record with the mode, the state, rho(1:nS, 1:nS) and the mode's SPF block
  (output.f90:254-275);
* ``density`` (``density``): per output time the time, then per DOF and
  state the reduced one-DOF density on the DOF's grid (Calc_Density,
  output.f90:332-445) and, for DVR DOFs, the populations of the basis
  functions underlying the DVR, which also go to the ``output`` block
  (output.f90:277-307);
* ``wpOvl`` (``wp_ovl``): the overlap matrix of the propagated wave packets
  (output.f90:309-322, Calc_SWP gmctdh_objects.f90:300-320); one packet
  here, so the 1 x 1 matrix <Psi|Psi>;
* ``log``: the tables read_input writes first (input.f90:391-413: state, DOF
  and mode counts, SPF structure, primitive basis representations, the
  Hamiltonian file).  The later tables of Macaw's log (operator bases,
  Hamiltonian terms, initial wavefunction) are not written.

Density-matrix convention: ``rho[s, m]`` as the engine returns it, the
transpose of Macaw's rho(j1, j2, m, s); Macaw reads only its upper triangle
(j1 <= j2) and so does this module.
"""
from __future__ import annotations

import struct

import numpy as np

from .restart import _write_record

AU2FS = 0.02418884326505


def _mode_block(prob, psi, m: int, s: int):
    """(SPF count, points per SPF, the SPF block of mode m for state s)."""
    ip, _ = prob.ipsiSPF()
    jel = s if prob.MultiSet[m] else 0
    n = int(prob.nSPF[m, s])
    npts = int(prob.npar()[m]) if prob.gwp_mode(m) else int(prob.ngpm()[m])
    start = int(ip[m, jel]) - 1
    return n, npts, psi[start:start + n * npts]


def _macaw_rho(rho, s: int, m: int, n: int) -> np.ndarray:
    """Macaw's rho(j1, j2) for j1 <= j2, completed as a Hermitian matrix from
    that upper triangle (the only part the reference reads)."""
    r = np.asarray(rho[s, m][:n, :n]).T
    up = np.triu(r)
    return up + np.triu(r, 1).conj().T


def densmat_records(prob, t: float, rho, psi, modes) -> list:
    """Records of one output time of the ``densmat`` file (output.f90:254-275):
    the time, then per listed mode (1-based) and state: mode, state (int64),
    rho(1:nS, 1:nS) in column order (complex128) and the SPF block."""
    recs = [struct.pack("<d", float(t))]
    for im in modes:
        m = int(im) - 1
        for s in range(prob.nstates):
            n, _, blk = _mode_block(prob, psi, m, s)
            r = np.asarray(rho[s, m][:n, :n]).T        # rho(j1, j2)
            recs.append(struct.pack("<qq", m + 1, s + 1) + np.asarray(r, "<c16").tobytes(order="F")
                        + np.ascontiguousarray(blk, "<c16").tobytes())
    return recs


def write_densmat(fh, prob, t: float, rho, psi, modes) -> None:
    for rec in densmat_records(prob, t, rho, psi, modes):
        _write_record(fh, rec)


def _dof_position(prob, dof: int):
    """(mode, position of the DOF inside the mode), both 0-based."""
    for m in range(prob.nmodes):
        for k in range(int(prob.nmodeDOF[m])):
            if int(prob.iDOF[k, m]) == dof + 1:
                return m, k
    raise ValueError(f"DOF {dof + 1} belongs to no mode")


def _gau_ovl_1d(a1: float, a2: float, x1: complex, x2: complex) -> complex:
    """<g1|g2> of normalised 1-D Gaussians exp(a x^2 + xi x), a < 0
    (GauOvl1D, gmctdh_objects.f90:1476-1493)."""
    c = x1.real ** 2 / a1 + x2.real ** 2 / a2 - (np.conj(x1) + x2) ** 2 / (a1 + a2)
    return np.exp(0.25 * c) * np.sqrt(4.0 * a1 * a2 / (a1 + a2) ** 2)


def calc_density(prob, rho, psi, dof: int, trafo=None):
    """Reduced density of one DOF (0-based) for every state (Calc_Density,
    output.f90:332-445): (Density[nG, nstates], DensityPop[nG, nstates] or
    None for a Gaussian DOF).

    DVR DOF: the density matrix on the DOF's grid,
    DM(g1, g2) = sum_{j1 j2} rho(j1, j2) sum_r conj(phi_j1(r, g2)) phi_j2(r, g1)
    over the other DOFs r of the mode (rho completed from its upper
    triangle), Density = diag DM (grid-point populations: the SPFs hold DVR
    amplitudes) and DensityPop(g) = T(g, :) DM T(g, :)^T with T the
    DVR-to-FBR transformation (rows: basis functions).
    Gaussian DOF: the Gaussians' product densities on the equidistant grid
    of the DOF's declaration, times the grid step."""
    m, k = _dof_position(prob, dof)
    nG = int(prob.ngp[dof])
    ns = prob.nstates
    dens = np.zeros((nG, ns))
    if prob.gwp_mode(m):
        x = np.asarray(prob.grid[:nG, dof])
        delta = (x[-1] - x[0]) / (nG - 1)
        nmd = int(prob.nmodeDOF[m])
        pi24 = 1.0 / np.arcsin(1.0) ** 2
        for s in range(ns):
            n, _, blk = _mode_block(prob, psi, m, s)
            xi = blk.reshape(n, nmd).T                  # xi[d, j]
            a = prob.gwp_a[:nmd, :n, m]
            R = _macaw_rho(rho, s, m, n)
            for j1 in range(n):
                a1, xr1 = a[k, j1], xi[k, j1].real ** 2
                for j2 in range(n):
                    a2, xr2 = a[k, j2], xi[k, j2].real ** 2
                    c = R[j1, j2] * (pi24 * a1 * a2 * np.exp(xr1 / a1 + xr2 / a2)) ** 0.25
                    for kg in range(nmd):
                        if kg != k:
                            c = c * _gau_ovl_1d(a[kg, j1], a[kg, j2], xi[kg, j1], xi[kg, j2])
                    dens[:, s] += (c * np.exp(x ** 2 * (a1 + a2) + x * (np.conj(xi[k, j1]) + xi[k, j2]))).real
        return dens * delta, None
    if trafo is None:
        raise ValueError(f"DOF {dof + 1}: the DVR-to-FBR transformation is required")
    pops = np.zeros((nG, ns))
    sizes = [int(prob.ngp[int(prob.iDOF[q, m]) - 1]) for q in range(int(prob.nmodeDOF[m]))]
    ll = int(np.prod(sizes[:k], dtype=np.int64))
    nl = int(np.prod(sizes[k + 1:], dtype=np.int64))
    T = np.asarray(trafo)[:nG, :nG]
    for s in range(ns):
        n, npts, blk = _mode_block(prob, psi, m, s)
        # phi_j(r, g): SPF j on the mode grid, the DOF's grid index g, the
        # other DOFs r (MatricisePsi, hamilton.f90:664-699)
        Q = blk.reshape(n, npts).reshape(n, nl, nG, ll).transpose(0, 1, 3, 2).reshape(n, nl * ll, nG)
        R = _macaw_rho(rho, s, m, n)
        DM = np.einsum("ab,arh,brg->gh", R, Q.conj(), Q)   # DM(g1=g, g2=h)
        dens[:, s] = DM.diagonal().real
        pops[:, s] = np.einsum("ga,ab,gb->g", T, DM, T).real
    return dens, pops


def density_records(prob, t: float, rho, psi, trafos) -> tuple:
    """Records of one output time of the ``density`` file and the lines the
    reference adds to ``output`` (output.f90:277-307)."""
    recs = [struct.pack("<d", float(t))]
    lines = []
    dvr_dofs = [d for d in range(prob.ndof) if not prob.gwp_mode(_dof_position(prob, d)[0])]
    if dvr_dofs:
        lines.append(" Population of the basis set underlying the DVR")
    for d in range(prob.ndof):
        dens, pops = calc_density(prob, rho, psi, d, trafos.get(d))
        nG = dens.shape[0]
        if pops is None:
            for s in range(prob.nstates):
                recs.append(struct.pack("<q", d + 1) + np.asarray(dens[:, s], "<f8").tobytes())
            continue
        lines.append(f"   DOF:{d + 1:2d}")
        for s in range(prob.nstates):
            recs.append(struct.pack("<q", d + 1) + np.asarray(dens[:, s], "<f8").tobytes()
                        + np.asarray(pops[:, s], "<f8").tobytes())
            if nG == 2:
                lines.append(f"     State {s + 1:2d}:   {pops[0, s]:7.4f}  {pops[-1, s]:7.4f}")
            else:
                lines.append(f"     State {s + 1:2d}:   {pops[0, s]:7.4f}  {pops[1, s]:7.4f} ... "
                             f"{pops[-2, s]:7.4f}  {pops[-1, s]:7.4f}")
    return recs, lines


def wpovl_lines(t: float, overlap: complex, header: bool = False) -> list:
    """One output time of ``wpOvl`` for a single wave packet (output.f90:309-322)."""
    out = ["Wave packet overlap matrix"] if header else []
    out.append(f"Time: {t * AU2FS:7.2f} fs")
    # the formats (7x,6(8x,i3,8x)) and (i3,2x,6(2x,"(",f7.4,",",f7.4,")"))
    # with one item: format control stops at the next data edit descriptor,
    # so the first line ends at the i3 and the second after the literals
    # that precede the seventh f7.4
    out.append(" " * 15 + f"{1:3d}")
    out.append(f"{1:3d}  " + f"  ({overlap.real:7.4f},{overlap.imag:7.4f})  (")
    return out


def log_tables(prob, grids: dict, hamiltonian: str) -> list:
    """The first tables of Macaw's ``log`` (input.f90:391-413); `grids[d]` is
    the (kind, npts, par1, par2) declaration of DOF d (0-based)."""
    out = [f"Number of electronic states: {prob.nstates}", f"Number of degrees of freedom: {prob.ndof}",
           f"Number of logical modes: {prob.nmodes}", "", " SPF structure"]
    for m in range(prob.nmodes):
        dofs = "".join(f"{int(prob.iDOF[k, m])}  " for k in range(int(prob.nmodeDOF[m])))
        # the 2x after the last item of a record writes nothing
        spfs = "  ".join(f"{int(prob.nSPF[m, s])}" for s in range(prob.nstates))
        out.append(f"   Mode {m + 1} -     DOFs: {dofs} ;    SPFs: {spfs}")
    out += ["", " Primitive basis representations"]
    for d in range(prob.ndof):
        kind, npts, p1, p2 = grids[d]
        out.append(f"   DOF {d + 1} - {kind:<6s}   {npts}    {p1:10.4f}    {p2:10.4f}")
    # list-directed output of a character(len = 100) name (input.f90:24):
    # the blanks that pad it, and where a compiler breaks such a long
    # list-directed record, differ between compilers; written without them
    out += ["", " The Hamiltonian is read from file: " + hamiltonian]
    return out
