"""Ensembles of independent LVC wavepackets on one device.

Members are seeded LVC models with uneven shapes: 2-4 electronic states,
4-8 modes, 4-8 SPFs per mode and 16-32 HO-DVR points per mode, with
frequencies, intra-state couplings (kappa, gamma), inter-state couplings
(lambda) and state energies drawn from the same distributions as the
4-state 8-mode workload (``c2`` in :mod:`macaw_amd.lvc`), each member from its
own seed.  ``mg_ensemble_propagate`` (include/macaw_gpu.h) runs them with a
pool of host threads, every member on its own context and streams, largest
first, under a device-memory budget.

Command line:
  python -m macaw_amd.ensemble --members 256 --fs 1 --threads 8 --json out.json
runs the ensemble (each member's LDS brick size chosen by the autotuner
unless --no-autotune) in mixed precision and in complex FP64 and reports the
ensemble throughput (member-fs per wall second) and every member's
max |C_mixed(t) - C_FP64(t)| over the output times.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import tempfile
import time

import numpy as np

from .abi import MgOutput, f64p, i64p, make_mg_problem
from .lvc import LVCSpec, make_problem

AU2FS = 0.02418884326505


def member_spec(k: int, seed: int = 4000, tfinal: float = 1.0, eps_int: float = 1e-7) -> LVCSpec:
    """Shape and energies of ensemble member k (deterministic in (seed, k))."""
    rng = np.random.default_rng([seed, k])
    ns = int(rng.integers(2, 5))
    nm = int(rng.integers(4, 9))
    nspf = int(rng.integers(4, 9))
    npts = int(rng.integers(16, 33))
    gaps = tuple(float(g) for g in np.sort(rng.uniform(0.0, 0.9, size=ns)))
    npairs = ns * (ns - 1) // 2
    n_lambda = int(rng.integers(1, min(6, nm * npairs) + 1))
    return LVCSpec(f"e{k}", int(rng.integers(1, 2**31 - 1)), ns, nm, npts, nspf, tfinal,
                   omega=(0.05, 0.40), kappa_sd=0.12, gamma_sd=0.02, n_lambda=n_lambda,
                   lam=(0.05, 0.20), gaps=gaps, eps_int=eps_int, eps_psi=eps_int)


def members(n: int, seed: int = 4000, tfinal: float = 1.0, eps_int: float = 1e-7):
    """Problems (with initial states) of members 0..n-1."""
    out = []
    for k in range(n):
        with tempfile.TemporaryDirectory() as d:
            out.append(make_problem(member_spec(k, seed, tfinal, eps_int), d))
    return out


def propagate(probs, t_fs, precision=1, threads=8, mem_fraction=0.7, autotune=False, strict=True, progress=False):
    """Run mg_ensemble_propagate; returns (outs per member, pops per member,
    member wall seconds, total wall seconds).  strict=False: failed members
    are returned with no outputs, and the result gains (status, message)."""
    from .engine import EngineError, lib
    L = lib()
    n = len(probs)
    keep, descs, a_ptrs, p_ptrs = [], [], [], []
    for p in probs:
        mgp, k = make_mg_problem(p)
        a = np.ascontiguousarray(p.A0, np.complex128)
        s = np.ascontiguousarray(p.psi0, np.complex128)
        keep += [k, a, s, mgp]
        descs.append(C.pointer(mgp))
        a_ptrs.append(a.ctypes.data_as(f64p))
        p_ptrs.append(s.ctypes.data_as(f64p))
    wt = min(float(p.writetime) for p in probs)
    max_out = int(t_fs / AU2FS / wt + 0.5) + 2
    outs = (MgOutput * (n * max_out))()
    pops = np.zeros((n, max_out, 8))
    n_out = np.zeros(n, np.int64)
    status = np.zeros(n, np.int64)
    wall = np.zeros(n)
    err = C.create_string_buffer(512)
    PP = C.POINTER(type(descs[0].contents)) if n else C.c_void_p
    probs_arr = (PP * n)(*descs)
    a_arr = (f64p * n)(*a_ptrs)
    p_arr = (f64p * n)(*p_ptrs)
    t0 = time.perf_counter()
    flags = (1 if autotune else 0) | (2 if progress else 0)
    rc = L.mg_ensemble_propagate(probs_arr, a_arr, p_arr, n, precision, threads, mem_fraction, flags,
                                 0.0, t_fs / AU2FS, outs, pops.ctypes.data_as(f64p), max_out,
                                 n_out.ctypes.data_as(i64p), status.ctypes.data_as(i64p), wall.ctypes.data_as(f64p),
                                 err, len(err))
    total = time.perf_counter() - t0
    if rc != 0 and strict:
        raise EngineError(f"ensemble: {err.value.decode()} (status {status.tolist()})")
    ok = status == 0
    res_o = [list(outs[i * max_out: i * max_out + int(n_out[i])]) if ok[i] else [] for i in range(n)]
    res_p = [pops[i, : int(n_out[i]), : probs[i].nstates].copy() if ok[i] else None for i in range(n)]
    if strict:
        return res_o, res_p, wall, total
    return res_o, res_p, wall, total, status, err.value.decode()


def autocorr(outs):
    return np.array([complex(o.auto_re, o.auto_im) for o in outs])


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--seed", type=int, default=4000)
    ap.add_argument("--fs", type=float, default=1.0)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--mem-fraction", type=float, default=0.7)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default LDS brick size instead of timing the candidates per member")
    ap.add_argument("--only", choices=("mixed", "fp64"), help="run one precision only (no comparison)")
    ap.add_argument("--json")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    probs = members(a.members, a.seed, a.fs)
    sizes = [int(p.nAConf().sum()) for p in probs]
    log(f"[ensemble] {a.members} members built in {time.perf_counter() - t0:.1f} s; configurations per member "
        f"min {min(sizes)}, median {int(np.median(sizes))}, max {max(sizes)}, total {sum(sizes)}")
    res = {"members": a.members, "fs_per_member": a.fs, "threads": a.threads, "configurations": sizes}
    runs = {}
    for tag, prec in (("mixed", 1), ("fp64", 0)):
        if a.only and tag != a.only:
            continue
        log(f"[ensemble] {tag}: propagating {a.members} members over {a.fs} fs")
        outs, pops, wall, total, status, msg = propagate(probs, a.fs, prec, a.threads, a.mem_fraction, a.autotune,
                                                         strict=False, progress=True)
        runs[tag] = (outs, pops, status)
        nok = int((status == 0).sum())
        res[tag] = {"wall_s": total, "ensemble_fs_per_s": nok * a.fs / total, "member_wall_s": wall.tolist(),
                    "failed": [int(i) for i in np.nonzero(status)[0]], "first_error": msg}
        log(f"[ensemble] {tag}: {total:.1f} s, {nok * a.fs / total:.3f} member-fs/s "
            f"(slowest member {wall.max():.1f} s); {a.members - nok} failed{': ' + msg if msg else ''}")
    if a.only:
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(res, fh)
        return 0
    dC = []
    dP = []
    for i in range(a.members):
        if runs["mixed"][2][i] or runs["fp64"][2][i]:
            dC.append(None)
            dP.append(None)
            continue
        cm, cd = autocorr(runs["mixed"][0][i]), autocorr(runs["fp64"][0][i])
        m = min(len(cm), len(cd))
        dC.append(float(np.max(np.abs(cm[:m] - cd[:m]))))
        dP.append(float(np.max(np.abs(runs["mixed"][1][i][:m] - runs["fp64"][1][i][:m]))))
    res["max_dC_vs_fp64"] = dC
    res["max_dpop_vs_fp64"] = dP
    good = [x for x in dC if x is not None]
    if good:
        log(f"[ensemble] per-member max|dC| vs FP64 ({len(good)} members): median {np.median(good):.2e}, "
            f"max {max(good):.2e}; max|dP| {max(x for x in dP if x is not None):.2e}")
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(res, fh)
    return 0


if __name__ == "__main__":
    sys.exit(main())
