"""ctypes mirror of the C-ABI structs in include/macaw_gpu.h and helpers that
turn a :class:`~macaw_amd.problem.Problem` into an ``mg_problem``."""
from __future__ import annotations

import ctypes as C

import numpy as np

i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class MgProblem(C.Structure):
    _fields_ = [
        ("nstates", C.c_int64), ("nmodes", C.c_int64), ("ndof", C.c_int64),
        ("maxdof", C.c_int64),
        ("nmodeDOF", i64p), ("iDOF", i64p), ("MultiSet", i64p), ("nSPF", i64p),
        ("ngp", i64p), ("ngpmax", C.c_int64),
        ("grid", f64p), ("d1", f64p), ("d2", f64p), ("xd1", f64p),
        ("nham", C.c_int64), ("ae", f64p), ("lhs_el", i64p), ("rhs_el", i64p),
        ("iOpDof", i64p),
        ("eps_int", C.c_double), ("eps_psi", C.c_double), ("eps_rho", C.c_double),
        ("writetime", C.c_double), ("min_stepsize", C.c_double),
        ("int_type_psi", C.c_int64),
        ("dof_gwp", i64p), ("gwp_a", f64p), ("eps_C", C.c_double), ("eps_ovl", C.c_double),
        ("nopparam", C.c_int64), ("gridoper", f64p),
    ]


class MgOutput(C.Structure):
    _fields_ = [("time", C.c_double), ("auto_re", C.c_double), ("auto_im", C.c_double),
                ("norm", C.c_double), ("energy", C.c_double)]


class MgAttempt(C.Structure):
    _fields_ = [("time", C.c_double), ("h", C.c_double), ("err_A", C.c_double),
                ("err_psi", C.c_double), ("accepted", C.c_int64),
                ("nkrylov1", C.c_int64), ("nkrylov2", C.c_int64)]


def _i64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.int64))


def _f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def cptr(a: np.ndarray, typ=f64p):
    return a.ctypes.data_as(typ)


def make_mg_problem(prob):
    """Return (MgProblem, keepalive list).  Arrays keep Fortran order."""
    keep = [
        _i64(prob.nmodeDOF), _i64(prob.iDOF), _i64(prob.MultiSet), _i64(prob.nSPF),
        _i64(prob.ngp), _f64(prob.grid), _f64(prob.d1), _f64(prob.d2), _f64(prob.xd1),
        np.ascontiguousarray(prob.ae.astype(np.complex128)).view(np.float64),
        _i64(prob.lhs_el), _i64(prob.rhs_el), _i64(prob.iOpDof),
    ]
    (nmd, idof, ms, nspf, ngp, grid, d1, d2, xd1, ae, lhs, rhs, iop) = keep
    gwp = getattr(prob, "dof_gwp", None)
    if gwp is not None and np.any(gwp):
        dg, ga = _i64(gwp), _f64(prob.gwp_a)
        keep += [dg, ga]
        dgp, gap = cptr(dg, i64p), cptr(ga)
    else:
        dgp, gap = None, None
    gop = getattr(prob, "gridoper", None)
    if gop is not None and gop.size:
        go = _f64(gop)
        keep += [go]
        nop, gop_p = int(gop.shape[1]), cptr(go)
    else:
        nop, gop_p = 0, None
    p = MgProblem(
        nstates=prob.nstates, nmodes=prob.nmodes, ndof=prob.ndof, maxdof=prob.maxDOF,
        nmodeDOF=cptr(nmd, i64p), iDOF=cptr(idof, i64p), MultiSet=cptr(ms, i64p),
        nSPF=cptr(nspf, i64p), ngp=cptr(ngp, i64p), ngpmax=prob.ngpMax,
        grid=cptr(grid), d1=cptr(d1), d2=cptr(d2), xd1=cptr(xd1),
        nham=prob.nHam, ae=cptr(ae), lhs_el=cptr(lhs, i64p), rhs_el=cptr(rhs, i64p),
        iOpDof=cptr(iop, i64p),
        eps_int=prob.eps_int, eps_psi=prob.eps_psi, eps_rho=prob.eps_rho,
        writetime=prob.writetime, min_stepsize=prob.min_stepsize,
        int_type_psi=prob.int_type_psi,
        dof_gwp=dgp, gwp_a=gap,
        eps_C=getattr(prob, "eps_C", 1e-8), eps_ovl=getattr(prob, "eps_ovl", 1e-8),
        nopparam=nop, gridoper=gop_p)
    return p, keep


def cplx_ptr(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(f64p)
