"""Problem descriptor file for the Fortran driver (macaw_amd/fortran/drive_cmf2.f90).

Stream-unformatted layout, native endianness: 9 int64 (nstates, nmodes, ndof,
maxdof, ngpmax, nham, nconf, dimpsi, int_type_psi), 6 float64 (eps_int,
eps_psi, eps_rho, writetime, min_stepsize, tfinal), then the int64 tables
nmodeDOF, iDOF, MultiSet, nSPF, ngp, lhs_el, rhs_el, iOpDof and the float64 /
complex128 arrays grid, d1, d2, xd1, ae, A0, psi0, all in Fortran order.
"""
from __future__ import annotations

import numpy as np


def write_descriptor(prob, path: str, tfinal: float) -> None:
    f = lambda a, t: np.asarray(a, dtype=t).reshape(-1, order="F")
    hdr = np.array([prob.nstates, prob.nmodes, prob.ndof, prob.maxDOF, prob.ngpMax, prob.nHam,
                    prob.nAConftot, prob.dimpsi, prob.int_type_psi], dtype=np.int64)
    par = np.array([prob.eps_int, prob.eps_psi, prob.eps_rho, prob.writetime, prob.min_stepsize,
                    tfinal], dtype=np.float64)
    with open(path, "wb") as fh:
        for a in (hdr, par):
            fh.write(a.tobytes())
        for a in (prob.nmodeDOF, prob.iDOF, prob.MultiSet, prob.nSPF, prob.ngp, prob.lhs_el,
                  prob.rhs_el, prob.iOpDof):
            fh.write(f(a, np.int64).tobytes())
        for a in (prob.grid, prob.d1, prob.d2, prob.xd1):
            fh.write(f(a, np.float64).tobytes())
        for a in (prob.ae, prob.A0, prob.psi0):
            fh.write(np.ascontiguousarray(a, dtype=np.complex128).tobytes())


def read_auto(path: str) -> np.ndarray:
    """Columns of a Macaw `auto` file: time [fs], Re, Im, |C|."""
    return np.loadtxt(path, comments="#")
