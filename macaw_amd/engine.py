"""Python binding of the HIP engine's C-ABI (include/macaw_gpu.h).

The product path: every call goes to libmacaw_gpu.so (HIP, gfx950).  There is
no CPU fallback; loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import MgAttempt, MgOutput, MgProblem, cplx_ptr, f64p, i64p, make_mg_problem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmacaw_gpu.so")
_lib = None


class EngineError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(f"HIP engine not built: {LIB_PATH} missing (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mg_create.restype = C.c_int
        L.mg_create.argtypes = [C.POINTER(MgProblem), C.c_int64, C.POINTER(vp)]
        L.mg_destroy.argtypes = [vp]
        L.mg_last_error.restype = C.c_char_p
        L.mg_last_error.argtypes = [vp]
        L.mg_nconf.restype = C.c_int64
        L.mg_nconf.argtypes = [vp]
        L.mg_dimpsi.restype = C.c_int64
        L.mg_dimpsi.argtypes = [vp]
        L.mg_dims.argtypes = [vp, i64p, i64p]
        L.mg_stepsize.argtypes = [vp, f64p]
        L.mg_set_stepsize.argtypes = [vp, C.c_double]
        L.mg_apply_h_profile.argtypes = [vp, vp, vp, f64p, i64p]
        L.mg_set_replay.argtypes = [vp, C.POINTER(MgAttempt), C.c_int64]
        L.mg_profile_copy.argtypes = [vp, C.c_int64, f64p]
        L.mg_calc_derivs.argtypes = [vp, f64p, f64p]
        L.mg_autotune.argtypes = [vp, C.c_int64, i64p, f64p]
        L.mg_propagate_vmf.argtypes = [vp, C.c_int64, C.c_double, C.c_double, C.POINTER(MgOutput), f64p, C.c_int64,
                                       i64p, C.POINTER(MgAttempt), C.c_int64, i64p]
        for f in ("mg_set_state", "mg_get_state", "mg_set_start"):
            getattr(L, f).argtypes = [vp, f64p, f64p]
        L.mg_calc_h1mode.argtypes = [vp, f64p]
        L.mg_calc_rho.argtypes = [vp, f64p]
        L.mg_calc_meanfield.argtypes = [vp, f64p, f64p]
        L.mg_apply_h.argtypes = [vp, f64p, f64p]
        L.mg_apply_h_device.argtypes = [vp, vp, vp]
        L.mg_apply_h_stats.argtypes = [vp, f64p, f64p]
        L.mg_der_spf.argtypes = [vp, C.c_int64, f64p, f64p]
        L.mg_cmf2_step.argtypes = [vp, C.c_double, C.c_double, f64p, f64p]
        L.mg_observables.argtypes = [vp, C.c_double, C.POINTER(MgOutput), f64p]
        L.mg_propagate.argtypes = [vp, C.c_double, C.c_double, C.POINTER(MgOutput), f64p, C.c_int64,
                                   i64p, C.POINTER(MgAttempt), C.c_int64, i64p]
        L.mg_prepare_gwp.argtypes = [vp]
        L.mg_gwp_overlaps.argtypes = [vp, f64p, f64p]
        L.mg_der_spf_gwp.argtypes = [vp, C.c_int64, f64p, f64p, f64p, f64p]
        L.mg_a_to_b.argtypes = [vp, f64p, f64p, f64p]
        f32p = C.POINTER(C.c_float)
        L.mg_mlp_train.argtypes = [vp, f32p, f32p, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_uint64, f32p,
                                   f64p]
        L.mg_mlp_forward.argtypes = [vp, f32p, f32p, C.c_int64, f32p]
        L.mg_mlp_forward_noalloc.argtypes = [vp, f32p, f32p, C.c_int64, f32p]
        L.mg_mlp_grad.argtypes = [vp, f32p, f32p, f32p, C.c_int64, f32p, f64p]
        L.mg_set_step_model.argtypes = [vp, f32p]
        L.mg_log_step_features.argtypes = [vp, C.c_int64]
        L.mg_step_features.argtypes = [vp, f32p, f64p, C.c_int64, i64p]
        L.mg_controller_stats.argtypes = [vp, i64p, i64p]
        L.mg_set_graphs.argtypes = [vp, C.c_int64]
        L.mg_set_normalize.argtypes = [vp, C.c_int64]
        L.mg_ensemble_propagate.argtypes = [vp, vp, vp, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int64,
                                            C.c_double, C.c_double, C.POINTER(MgOutput), f64p, C.c_int64, i64p,
                                            i64p, f64p, C.c_char_p, C.c_int64]
        _lib = L
    return _lib


class Engine:
    """One engine context (one HIP stream) for a problem descriptor."""

    def __init__(self, prob, precision: int = 0):
        self.prob = prob
        self.mgp, self._keep = make_mg_problem(prob)
        h = C.c_void_p()
        rc = lib().mg_create(C.byref(self.mgp), precision, C.byref(h))
        self.h = h
        if rc != 0:
            msg = lib().mg_last_error(h).decode() if h else "mg_create failed"
            lib().mg_destroy(h)
            self.h = None
            raise EngineError(msg)
        self.nconf = lib().mg_nconf(self.h)
        self.dimpsi = lib().mg_dimpsi(self.h)
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(lib().mg_dims(self.h, C.byref(a), C.byref(b)))
        self.nspfmax, self.nopmax = a.value, b.value
        self.set_state(prob.A0, prob.psi0)
        self.set_start(prob.A0, prob.psi0)

    def close(self):
        if self.h:
            lib().mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise EngineError(lib().mg_last_error(self.h).decode())

    def set_state(self, A, psi):
        A = np.ascontiguousarray(A, np.complex128)
        psi = np.ascontiguousarray(psi, np.complex128)
        self._check(lib().mg_set_state(self.h, cplx_ptr(A), cplx_ptr(psi)))

    def set_start(self, A, psi):
        A = np.ascontiguousarray(A, np.complex128)
        psi = np.ascontiguousarray(psi, np.complex128)
        self._check(lib().mg_set_start(self.h, cplx_ptr(A), cplx_ptr(psi)))

    def get_state(self):
        A = np.zeros(self.nconf, np.complex128)
        psi = np.zeros(self.dimpsi, np.complex128)
        self._check(lib().mg_get_state(self.h, cplx_ptr(A), cplx_ptr(psi)))
        return A, psi

    def h1mode(self, nopmax=None):
        nopmax = self.nopmax
        out = np.zeros((self.prob.nmodes, nopmax, self.nspfmax, self.nspfmax), np.complex128)
        self._check(lib().mg_calc_h1mode(self.h, cplx_ptr(out)))
        return out

    def rho(self):
        out = np.zeros((self.prob.nstates, self.prob.nmodes, self.nspfmax, self.nspfmax), np.complex128)
        self._check(lib().mg_calc_rho(self.h, cplx_ptr(out)))
        return out

    def meanfield(self, nopmax=None):
        shape = (self.prob.nmodes, self.nopmax, self.nspfmax, self.nspfmax)
        mf = np.zeros(shape, np.complex128)
        rmf = np.zeros(shape, np.complex128)
        self._check(lib().mg_calc_meanfield(self.h, cplx_ptr(mf), cplx_ptr(rmf)))
        return mf, rmf

    def apply_h(self, B):
        B = np.ascontiguousarray(B, np.complex128)
        HB = np.zeros_like(B)
        self._check(lib().mg_apply_h(self.h, cplx_ptr(B), cplx_ptr(HB)))
        return HB

    def apply_h_device(self, dB_ptr: int, dHB_ptr: int):
        self._check(lib().mg_apply_h_device(self.h, C.c_void_p(dB_ptr), C.c_void_p(dHB_ptr)))

    def apply_h_profile(self, dB_ptr: int = 0, dHB_ptr: int = 0):
        """Event-timed H.A; null pointers use the buffers of the last apply_h."""
        ms, n = C.c_double(0), C.c_int64(0)
        self._check(lib().mg_apply_h_profile(self.h, C.c_void_p(dB_ptr or None), C.c_void_p(dHB_ptr or None),
                                             C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def profile_copy(self, n):
        """Event-timed plain copy of n work-type elements (bandwidth reference)."""
        ms = C.c_double(0)
        self._check(lib().mg_profile_copy(self.h, n, C.byref(ms)))
        return ms.value

    def set_replay(self, attempts):
        n = len(attempts)
        arr = (MgAttempt * max(1, n))(*attempts)
        self._check(lib().mg_set_replay(self.h, arr, n))

    def stepsize(self):
        h = C.c_double(0)
        self._check(lib().mg_stepsize(self.h, C.byref(h)))
        return h.value

    def set_stepsize(self, h):
        """First attempted step of the next propagate call (a.u.)."""
        self._check(lib().mg_set_stepsize(self.h, float(h)))

    def apply_h_stats(self):
        b, f = C.c_double(0), C.c_double(0)
        self._check(lib().mg_apply_h_stats(self.h, C.byref(b), C.byref(f)))
        return b.value, f.value

    def der_spf(self, im, psi):
        psi = np.ascontiguousarray(psi, np.complex128)
        d = np.zeros_like(psi)
        self._check(lib().mg_der_spf(self.h, im, cplx_ptr(psi), cplx_ptr(d)))
        return d

    def cmf2_step(self, t, htry):
        hd, hn = C.c_double(0), C.c_double(0)
        self._check(lib().mg_cmf2_step(self.h, t, htry, C.byref(hd), C.byref(hn)))
        return hd.value, hn.value

    def observables(self, t):
        out = MgOutput()
        pops = np.zeros(self.prob.nstates)
        self._check(lib().mg_observables(self.h, t, C.byref(out), pops.ctypes.data_as(f64p)))
        return out, pops

    def autotune(self, reps=2):
        """Pick the LDS brick size (mode grouping) by timing H.A and the mean fields."""
        b, ms = C.c_int64(0), C.c_double(0)
        self._check(lib().mg_autotune(self.h, reps, C.byref(b), C.byref(ms)))
        return b.value, ms.value

    def calc_derivs(self):
        """VMF derivatives of the resident state: (-i H A, dpsi)."""
        dA = np.zeros(self.nconf, np.complex128)
        dpsi = np.zeros(self.dimpsi, np.complex128)
        self._check(lib().mg_calc_derivs(self.h, cplx_ptr(dA), cplx_ptr(dpsi)))
        return dA, dpsi

    def propagate_vmf(self, method, t0, tfinal, max_out=100000, max_att=200000):
        outs = (MgOutput * max_out)()
        pops = np.zeros((max_out, self.prob.nstates))
        atts = (MgAttempt * max_att)()
        no, na = C.c_int64(0), C.c_int64(0)
        self._check(lib().mg_propagate_vmf(self.h, method, t0, tfinal, outs, pops.ctypes.data_as(f64p), max_out,
                                           C.byref(no), atts, max_att, C.byref(na)))
        return list(outs[: no.value]), pops[: no.value], list(atts[: na.value])

    def propagate(self, t0, tfinal, max_out=100000, max_att=200000):
        outs = (MgOutput * max_out)()
        pops = np.zeros((max_out, self.prob.nstates))
        atts = (MgAttempt * max_att)()
        no, na = C.c_int64(0), C.c_int64(0)
        self._check(lib().mg_propagate(self.h, t0, tfinal, outs, pops.ctypes.data_as(f64p), max_out,
                                       C.byref(no), atts, max_att, C.byref(na)))
        return list(outs[: no.value]), pops[: no.value], list(atts[: na.value])

    # ---- Gaussian (G-MCTDH) modes ----
    def prepare_gwp(self):
        """S00, h1mode, D0 and D^H h D of the resident state (apply_h then
        works in the orthonormalised basis)."""
        self._check(lib().mg_prepare_gwp(self.h))

    def gwp_overlaps(self):
        ns, nm, n = self.prob.nstates, self.prob.nmodes, self.nspfmax
        S = np.zeros((ns, ns, nm, n, n), np.complex128)
        D = np.zeros((ns, nm, n, n), np.complex128)
        self._check(lib().mg_gwp_overlaps(self.h, cplx_ptr(S), cplx_ptr(D)))
        return S, D

    def der_spf_gwp(self, im, psi, D):
        psi = np.ascontiguousarray(psi, np.complex128)
        D = np.ascontiguousarray(D, np.complex128)
        dpsi = np.zeros_like(psi)
        dD = np.zeros_like(D)
        self._check(lib().mg_der_spf_gwp(self.h, im, cplx_ptr(psi), cplx_ptr(D), cplx_ptr(dpsi), cplx_ptr(dD)))
        return dpsi, dD

    def a_to_b(self, A, D):
        A = np.ascontiguousarray(A, np.complex128)
        D = np.ascontiguousarray(D, np.complex128)
        B = np.zeros_like(A)
        self._check(lib().mg_a_to_b(self.h, cplx_ptr(A), cplx_ptr(D), cplx_ptr(B)))
        return B

    def set_normalize(self, mode: int):
        """inorm_psi of propagate_wf: 0 never, 1 first interval, 2 every
        output interval, 3 every step (renormalise A by Trace_wf)."""
        self._check(lib().mg_set_normalize(self.h, int(mode)))

    def set_graphs(self, enable=True):
        """Replay the cached launch plans as HIP graphs (default) or launch kernel by kernel."""
        self._check(lib().mg_set_graphs(self.h, 1 if enable else 0))

    # ---- learned step controller ----
    def mlp_train(self, X, t, epochs=30, batch=256, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        params = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double(0)
        self._check(lib().mg_mlp_train(self.h, _f32(X), _f32(t), len(t), epochs, batch, lr, seed, _f32(params),
                                       C.byref(loss)))
        return params, loss.value

    def mlp_forward(self, params, X):
        X = np.ascontiguousarray(X, np.float32)
        params = np.ascontiguousarray(params, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        self._check(lib().mg_mlp_forward(self.h, _f32(params), _f32(X), X.shape[0], _f32(y)))
        return y

    def mlp_forward_noalloc(self, params, X):
        """mg_mlp_forward without device allocation (at most
        MG_MLP_NOALLOC_MAX_ROWS = 4096 rows)."""
        X = np.ascontiguousarray(X, np.float32)
        params = np.ascontiguousarray(params, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        self._check(lib().mg_mlp_forward_noalloc(self.h, _f32(params), _f32(X), X.shape[0], _f32(y)))
        return y

    def mlp_grad(self, params, X, t):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        params = np.ascontiguousarray(params, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double(0)
        self._check(lib().mg_mlp_grad(self.h, _f32(params), _f32(X), _f32(t), len(t), _f32(g), C.byref(loss)))
        return g, loss.value

    def set_step_model(self, params):
        if params is None:
            self._check(lib().mg_set_step_model(self.h, None))
            return
        self._model = np.ascontiguousarray(params, np.float32)
        self._check(lib().mg_set_step_model(self.h, _f32(self._model)))

    def log_step_features(self, enable=True):
        self._check(lib().mg_log_step_features(self.h, 1 if enable else 0))

    def step_features(self, max_rows=1000000):
        f = np.zeros((max_rows, MLP_NFEAT), np.float32)
        h = np.zeros(max_rows)
        n = C.c_int64(0)
        self._check(lib().mg_step_features(self.h, _f32(f), h.ctypes.data_as(f64p), max_rows, C.byref(n)))
        return f[: n.value].copy(), h[: n.value].copy()

    def controller_stats(self):
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(lib().mg_controller_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


MLP_NFEAT = 12
MLP_NPARAM = 5081


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))

