"""Throughput of the CMF2 propagation on one MI355X.

A "step" is one output interval (tout = 0.5 fs) of CMF2 propagation of the
4-state 8-mode LVC model (HO-DVR 32 points, 8 SPFs per mode and state, 8^8
configurations per state), including the observables written at the start of
the interval, exactly as Macaw's propagate_wf does it.  W intervals (which
include the reference's step-size ramp from tout*eps_int) run untimed, then K
intervals are timed.  value = femtoseconds propagated per wall second.

Also reported: the roofline of the dominant kernel (the grouped mode-product
pass of H.A) timed with a HIP event pair around every launch, and a CPU
baseline from the FP64 oracle run on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

AU2FS = 0.02418884326505
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
METRIC = "fs propagated/s at ≤1e-6 C(t) error vs FP64; mixed-prec vs FP64, 1 GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_problem(name):
    from macaw_amd.lvc import make_problem
    with tempfile.TemporaryDirectory() as d:
        return make_problem(name, d)


def roofline(eng, prob, reps):
    # input resident in HBM: one host upload through apply_h, then event-timed
    # H.A on the engine's own device buffers
    n = prob.nAConftot
    rng = np.random.default_rng(1)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(n)
    eng.h1mode()
    eng.apply_h(x)
    eng.apply_h_profile()      # plan build + warm
    tot_ms, tot_n = 0.0, 0
    for _ in range(reps):
        ms, nl = eng.apply_h_profile()
        tot_ms += ms
        tot_n += nl
    bytes_alg, flops_alg = eng.apply_h_stats()
    avg_launch_ms = tot_ms / tot_n
    launches_per_ha = tot_n / reps
    flops_per_launch = flops_alg / launches_per_ha
    achieved = flops_per_launch / (avg_launch_ms * 1e-3) / 1e12
    return {
        "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
        "kernel": "k_group_run (grouped mode-product pass of H.A)",
        "avg_launch_ms": avg_launch_ms, "launches_per_HA": launches_per_ha,
        "HA_ms": avg_launch_ms * launches_per_ha, "HA_flops_alg": flops_alg, "HA_bytes_alg": bytes_alg,
        "HA_alg_GBs": bytes_alg / (avg_launch_ms * launches_per_ha * 1e-3) / 1e9,
    }


def cpu_baseline(prob, ha_per_fs, threads):
    """Oracle Calc_DHD_B at the full shape on `threads` host cores; the CPU
    fs/s is bounded above by 1 / (H.A applications per fs * t_HA)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(prob)
    o.h1mode()
    rng = np.random.default_rng(2)
    B = (rng.normal(size=prob.nAConftot) + 1j * rng.normal(size=prob.nAConftot)).astype(np.complex128)
    t0 = time.perf_counter()
    o.apply_h(B)
    t_ha = time.perf_counter() - t0
    return {
        "value": 1.0 / (ha_per_fs * t_ha), "unit": "fs/s", "cores": threads, "kind": "port",
        "sample": (f"one oracle Calc_DHD_B (H.A) at the full 4x8^8 shape took {t_ha:.2f} s on {threads} "
                   f"threads; CPU fs/s = 1/(H.A per fs x t_HA) with {ha_per_fs:.1f} H.A per fs counted "
                   "in the GPU run (Lanczos Krylov sizes + energies); mean fields and SPF work are "
                   "excluded, so this is an upper bound on the CPU rate"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--roofline-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU figures are not produced "
              "(see the sharding design in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from macaw_amd.engine import Engine

    log(f"building workload {args.workload}")
    prob = build_problem(args.workload)
    tout = prob.writetime
    log(f"nconf={prob.nAConftot} dimpsi={prob.dimpsi} nham={prob.nHam}")
    eng = Engine(prob)
    log(f"warmup: {args.warmup} output intervals")
    t_w = time.perf_counter()
    outs_w, _, atts_w = eng.propagate(0.0, (args.warmup - 1) * tout)
    log(f"warmup done in {time.perf_counter() - t_w:.1f} s, {len(atts_w)} attempts")
    t_a = args.warmup * tout
    log(f"timed: {args.steps} output intervals from {t_a * AU2FS:.2f} fs")
    t0 = time.perf_counter()
    outs, pops, atts = eng.propagate(t_a, t_a + (args.steps - 1) * tout)
    wall = time.perf_counter() - t0
    fs = args.steps * tout * AU2FS
    value = fs / wall
    nha = sum(a.nkrylov1 + a.nkrylov2 for a in atts) + len(outs)
    log(f"timed done: {wall:.2f} s for {fs:.2f} fs, {len(atts)} attempts "
        f"({sum(1 for a in atts if not a.accepted)} rejected), {nha} H.A")
    rl = roofline(eng, prob, args.roofline_reps)
    log(f"roofline: {rl['achieved']:.2f} TFLOP/s, H.A {rl['HA_ms']:.2f} ms")
    last = outs[-1]
    result = {
        "metric": METRIC, "value": value, "unit": "fs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": "LVC 4 states x 8 modes, HO-DVR 32, 8 SPFs (8^8 conf/state), CMF2 eps 1e-7",
                   "precision": "complex FP64 (mixed-precision form not built yet)",
                   "step": "one 0.5 fs output interval", "attempts": len(atts),
                   "HA_per_fs": nha / fs, "C_t_last": [last.auto_re, last.auto_im],
                   "norm_last": last.norm},
        "roofline": {k: rl[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": {k: v for k, v in rl.items() if k not in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
    }
    eng.close()
    if not args.no_cpu:
        log(f"cpu baseline on {args.cpu_threads} threads")
        result["cpu_baseline"] = cpu_baseline(prob, nha / fs, args.cpu_threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
