"""Throughput of the CMF2 propagation on one MI355X, mixed precision vs FP64.

A "step" is one output interval (tout = 0.5 fs) of CMF2 propagation of the
4-state 8-mode LVC model (HO-DVR 32 points, 8 SPFs per mode and state, 8^8
configurations per state), including the observables written at the start of
the interval, exactly as Macaw's propagate_wf does it.  W intervals (which
include the reference's step-size ramp from tout*eps_int) run untimed, then K
intervals are timed.  value = femtoseconds propagated per wall second by the
mixed-precision engine (complex FP32 H.A / Krylov / mean-field chains, FP64
state, SPFs and reductions).  The same W+K intervals are then run by the FP64
engine; max|dC(t)| over every output time is the accuracy gate (<= 1e-6) and
the FP64 rate is reported beside it.  Both runs take their step proposals
from the learned CMF2 controller (macaw_amd/step_controller.npy; the CMF2
error test still accepts or rejects every attempt) unless --controller
classical; a third run with the classical controller is reported beside them.

Also reported: the roofline of the dominant kernel (the grouped mode-product
pass of H.A) timed with a HIP event pair around every launch on the engine's
stream (algorithmic flops counted as 6 real flops per complex multiply-add,
the 8-flop figure beside it; HBM traffic per H.A application from the
committed counter runs), and CPU baselines from the FP64 oracle run on the
host: one complete CMF2 attempt at the full shape scaled by the attempts per
fs of the GPU run, and the first C1 femtoseconds end to end beside the FP64
engine on the same span.
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
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
METRIC = "fs propagated/s at ≤1e-6 C(t) error vs FP64; mixed-prec vs FP64, 1 GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def describe(name, prob):
    """Workload line of the JSON config, from the problem itself."""
    if name == "c2":
        return "LVC 4 states x 8 modes, HO-DVR 32, 8 SPFs (8^8 conf/state), CMF2 eps 1e-7"
    if name == "c3j":
        return ("LVC 3 states x 6 modes, 3 Gaussian modes (5 Gaussians, centres N(0, 0.2) from a restart file), "
                f"3 HO-DVR 28 modes (7 SPFs), {prob.nAConftot} configurations, CMF2 eps {prob.eps_int:g}, "
                "0.01 fs output intervals")
    gm = [m for m in range(prob.nmodes) if prob.gwp_mode(m)]
    kinds = (f"{len(gm)} Gaussian modes ({int(prob.nSPF[gm[0], 0])} Gaussians), " if gm else "")
    grid = [m for m in range(prob.nmodes) if not prob.gwp_mode(m)]
    npts = int(prob.ngp[int(prob.iDOF[0, grid[0]]) - 1]) if grid else 0
    return (f"LVC {prob.nstates} states x {prob.nmodes} modes, {kinds}{len(grid)} HO-DVR {npts} modes "
            f"({int(prob.nSPF[grid[0], 0]) if grid else 0} SPFs), {prob.nAConftot} configurations, "
            f"CMF2 eps {prob.eps_int:g}")


def build_problem(name):
    from macaw_amd.lvc import make_problem
    with tempfile.TemporaryDirectory() as d:
        prob = make_problem(name, d)
    if name == "c3j":
        # C3 with jittered Gaussian centres: the nearly coincident Gaussians
        # make the Gaussian-parameter equations stiff and eps_int = 1e-7 sits
        # below the SPF error estimate's rounding floor, so it runs at
        # eps_int 1e-5 and the bench windows are 0.01 fs output intervals
        prob.eps_int = 1e-5
        prob.writetime = 0.01 / AU2FS
    return prob


def roofline(eng, prob, reps, peak):
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
    # algorithmic flops: the engine counts 8 real flops per complex
    # multiply-add (4M); the roofline uses 6 (3M, the c64 count of the
    # workload's algorithmic work), the 4M figure is reported beside it
    flops_3m = flops_alg * 6.0 / 8.0
    achieved = flops_3m / launches_per_ha / (avg_launch_ms * 1e-3) / 1e12
    achieved_4m = flops_per_launch / (avg_launch_ms * 1e-3) / 1e12
    return {
        "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
        "frac": achieved / peak, "traffic": None,
        "kernel": "k_brick (mode-group brick programs of the H.A schedule)",
        "avg_launch_ms": avg_launch_ms, "launches_per_HA": launches_per_ha,
        "HA_ms": avg_launch_ms * launches_per_ha, "HA_flops_alg_3M": flops_3m, "HA_flops_alg_4M": flops_alg,
        "achieved_4M_TFLOPs": achieved_4m, "frac_4M": achieved_4m / peak, "HA_bytes_alg": bytes_alg,
        "HA_alg_GBs": bytes_alg / (avg_launch_ms * launches_per_ha * 1e-3) / 1e9,
    }


def gram_profile(eng, prob, peak, reps=5):
    """Density matrices and mean fields of the resident state: host-timed
    calls (each ends in a stream synchronisation and a copy of the small
    matrices), against the algorithmic flops of their all-but-one-mode
    contractions: rho, every (state, mode); mean fields, the Grams of the
    inter-state terms (mirror pairs counted once) on every mode."""
    nconf = prob.nAConf()
    n = prob.nSPF
    rho_flops = sum(8.0 * nconf[s] * n[m, s] for s in range(prob.nstates) for m in range(prob.nmodes))
    lhs, rhs = np.asarray(prob.lhs_el).ravel() - 1, np.asarray(prob.rhs_el).ravel() - 1
    inter = [t for t in range(prob.nHam) if lhs[t] != rhs[t]]
    mf_flops = 0.5 * sum(8.0 * nconf[rhs[t]] * n[m, rhs[t]] for t in inter for m in range(prob.nmodes))
    eng.rho()
    eng.meanfield()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.rho()
    t_rho = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.meanfield()
    t_mf = (time.perf_counter() - t0) / reps
    res = {"rho_ms": t_rho * 1e3, "rho_flops_alg": rho_flops, "rho_TFLOPs": rho_flops / t_rho / 1e12,
           "rho_frac": rho_flops / t_rho / 1e12 / peak, "meanfield_ms": t_mf * 1e3,
           "meanfield_gram_flops_alg": mf_flops, "meanfield_TFLOPs": mf_flops / t_mf / 1e12,
           "meanfield_frac": mf_flops / t_mf / 1e12 / peak}
    return {k: float(v) for k, v in res.items()}


def cpu_baseline(prob, att_per_fs, h_step, threads):
    """One complete CMF2 attempt of the FP64 oracle at the full shape on
    `threads` host cores (rho, h1mode and mean fields twice, two Lanczos
    builds and the final combination, four SPF half-step propagations,
    DeltaPsi: cmf2_base.f90:47-206), from the initial state with the median
    step of the GPU run; CPU fs/s = 1 / (CMF2 attempts per fs of the GPU run
    x t_attempt)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(prob)
    t0 = time.perf_counter()
    rc, hd, hn, atts = o.cmf2_step(0.0, h_step)
    t_att = (time.perf_counter() - t0) / max(1, len(atts))
    nk = [(a.nkrylov1, a.nkrylov2) for a in atts]
    del o
    return {
        "value": 1.0 / (att_per_fs * t_att), "unit": "fs/s", "cores": threads, "kind": "port",
        "sample": (f"one complete oracle CMF2 attempt at the full 4x8^8 shape (rho, h1mode and mean fields "
                   f"twice, Lanczos builds of {nk[0][0]}+{nk[0][1]} Krylov vectors, 4 SPF half-step "
                   f"propagations, DeltaPsi) took {t_att:.1f} s on {threads} threads with the GPU run's "
                   f"median step {h_step * AU2FS:.4f} fs; CPU fs/s = 1/(attempts per fs x t_attempt) with "
                   f"{att_per_fs:.1f} attempts per fs counted in the GPU run"),
    }


def c1_baseline(span_fs, threads, warm=True):
    """End-to-end C1 (2 states x 5 modes, HO-DVR 24, 6 SPFs; propagate_wf with
    CMF2 from t = 0): the FP64 oracle on `threads` host cores and the FP64
    engine on the GPU over the same span, with the same arithmetic (each
    run free, its own controller)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from macaw_amd.engine import Engine
    prob = build_problem("c1")
    tf = span_fs / AU2FS
    o = Oracle(prob)
    t0 = time.perf_counter()
    oo, _, oa = o.propagate(0.0, tf)
    t_cpu = time.perf_counter() - t0
    e = Engine(prob, precision=0)
    if warm:
        e.propagate(0.0, 0.0)
        e.set_state(prob.A0, prob.psi0)
    t0 = time.perf_counter()
    eo, _, ea = e.propagate(0.0, tf)
    t_gpu = time.perf_counter() - t0
    e.close()
    dc = max(abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im)) for a, b in zip(eo, oo))
    return {"workload": "c1: LVC 2 states x 5 modes, HO-DVR 24, 6 SPFs, CMF2 eps 1e-7", "span_fs": span_fs,
            "cpu_fs_per_s": span_fs / t_cpu, "cpu_wall_s": t_cpu, "cpu_threads": threads,
            "cpu_attempts": len(oa), "gpu_fp64_fs_per_s": span_fs / t_gpu, "gpu_wall_s": t_gpu,
            "gpu_attempts": len(ea), "max_abs_dC_gpu_vs_cpu": dc}


def measured_traffic(workload, precision, launches_per_ha):
    """HBM bytes per launch of the dominant kernel from the committed counter
    runs (FETCH_SIZE / WRITE_SIZE), used only while the launch schedule they
    were measured on is unchanged."""
    path = os.path.join(ROOT, "profiles", "k_brick_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if (t.get("workload"), t.get("precision"), t.get("launches_per_HA")) != (workload, precision, launches_per_ha):
        return None
    if "bytes_per_HA" not in t:
        return None
    return t


def step_model(kind, path=None):
    """Network of the learned CMF2 step controller, or None (classical)."""
    if kind == "classical":
        return None
    import numpy as np
    return np.load(path or os.path.join(ROOT, "macaw_amd", "step_controller.npy"))


def run_engine(prob, precision, warmup, steps, autotune=True, model=None, tag=None):
    from macaw_amd.engine import Engine
    tout = prob.writetime
    eng = Engine(prob, precision=precision)
    if model is not None:
        eng.set_step_model(model)
    tag = tag or ("mixed" if precision else "fp64")
    brick = None
    if autotune:
        t_a = time.perf_counter()
        brick, score = eng.autotune()
        log(f"[{tag}] autotune: brick {brick} (score {score:.2f} ms) in {time.perf_counter() - t_a:.1f} s")
    log(f"[{tag}] warmup: {warmup} output intervals")
    t_w = time.perf_counter()
    outs_w, _, atts_w = eng.propagate(0.0, (warmup - 1) * tout)
    log(f"[{tag}] warmup done in {time.perf_counter() - t_w:.1f} s, {len(atts_w)} attempts")
    t_a = warmup * tout
    log(f"[{tag}] timed: {steps} output intervals from {t_a * AU2FS:.2f} fs")
    t0 = time.perf_counter()
    outs, pops, atts = eng.propagate(t_a, t_a + (steps - 1) * tout)
    wall = time.perf_counter() - t0
    fs = steps * tout * AU2FS
    nha = sum(a.nkrylov1 + a.nkrylov2 for a in atts) + len(outs)
    log(f"[{tag}] timed done: {wall:.2f} s for {fs:.2f} fs ({fs / wall:.4f} fs/s), {len(atts)} attempts "
        f"({sum(1 for a in atts if not a.accepted)} rejected), {nha} H.A")
    return eng, {"outs": outs_w + outs, "wall": wall, "fs": fs, "atts": atts, "nha": nha, "brick": brick}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--roofline-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 reference run (profiling)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--ha-only", action="store_true",
                    help="only the event-timed H.A applications and a plain copy of nconf elements "
                         "(the workload of the counter passes)")
    ap.add_argument("--precision", type=int, default=1, help="--ha-only: 1 mixed (default), 0 FP64")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default LDS brick size")
    ap.add_argument("--controller", choices=("learned", "classical"), default="learned",
                    help="CMF2 step proposal of every run (the CMF2 error test accepts or rejects either way)")
    ap.add_argument("--step-model", help="network file of the learned controller (default: the shipped one)")
    ap.add_argument("--no-classical", action="store_true",
                    help="skip the extra mixed run with the classical controller")
    ap.add_argument("--c1-fs", type=float, default=5.0,
                    help="span of the end-to-end C1 baseline (oracle on the host vs FP64 engine; 0 skips, "
                         "100 is the whole C1 workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU figures are not produced "
              "(a multi-GPU engine would shard the wavefunction by electronic state)", file=sys.stderr)
        sys.exit(2)

    log(f"building workload {args.workload}")
    prob = build_problem(args.workload)
    log(f"nconf={prob.nAConftot} dimpsi={prob.dimpsi} nham={prob.nHam}")
    if args.ha_only:
        from macaw_amd.engine import Engine
        eng = Engine(prob, precision=args.precision)
        rl = roofline(eng, prob, args.roofline_reps, PEAK_FP32_TFLOPS if args.precision else PEAK_FP64_TFLOPS)
        es = 8 if args.precision else 16
        cms = min(eng.profile_copy(prob.nAConftot) for _ in range(3))
        rl["copy_ms"] = cms
        rl["copy_bytes"] = 2.0 * es * prob.nAConftot
        rl["copy_GBs"] = rl["copy_bytes"] / (cms * 1e-3) / 1e9
        eng.close()
        print(json.dumps({"ha_only": True, "precision": args.precision, **rl}), flush=True)
        return
    model = step_model(args.controller, args.step_model)
    eng, mx = run_engine(prob, 1, args.warmup, args.steps, not args.no_autotune, model)
    rl = roofline(eng, prob, args.roofline_reps, PEAK_FP32_TFLOPS)
    log(f"roofline (mixed): {rl['achieved']:.2f} TFLOP/s, H.A {rl['HA_ms']:.2f} ms")
    gp = gram_profile(eng, prob, PEAK_FP32_TFLOPS)
    log(f"rho {gp['rho_ms']:.2f} ms ({gp['rho_TFLOPs']:.1f} TFLOP/s), mean fields {gp['meanfield_ms']:.2f} ms "
        f"({gp['meanfield_TFLOPs']:.1f} TFLOP/s, Gram contractions)")
    rl.update({"contractions_" + k: v for k, v in gp.items()})
    eng.close()
    value = mx["fs"] / mx["wall"]
    last = mx["outs"][-1]
    result = {
        "metric": METRIC, "value": value, "unit": "fs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": mx["wall"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64/c128", "data": "synthetic",
        "config": {"workload": describe(args.workload, prob),
                   "precision": "mixed: complex FP32 H.A, Krylov v1.., mean-field chains; FP64 A, SPFs, "
                                "rho, mean fields, reductions, DeltaPsi",
                   "step": "one 0.5 fs output interval", "attempts": len(mx["atts"]),
                   "rejected": sum(1 for a in mx["atts"] if not a.accepted),
                   "controller": args.controller,
                   "HA_per_fs": mx["nha"] / mx["fs"], "lds_brick": mx["brick"],
                   "C_t_last": [last.auto_re, last.auto_im],
                   "norm_last": last.norm},
        "roofline": {k: rl[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": {k: v for k, v in rl.items() if k not in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
    }
    result["roofline_detail"]["peak_note"] = "MI355X dense FP32 (vector and matrix) 157.3 TFLOP/s"
    tr = measured_traffic(args.workload, 1, rl["launches_per_HA"])
    if tr is not None:
        # HBM bytes of one whole H.A application (all its launches), beside
        # the algorithmic bytes of one H.A (read B once, write HB once)
        result["roofline"]["traffic"] = tr["bytes_per_HA"]
        result["roofline_detail"]["traffic_unit"] = "bytes per H.A application (all launches)"
        result["roofline_detail"]["traffic_per_launch"] = tr["bytes_per_HA"] / rl["launches_per_HA"]
        result["roofline_detail"]["traffic_over_alg"] = tr["bytes_per_HA"] / rl["HA_bytes_alg"]
        result["roofline_detail"]["traffic_source"] = tr["source"]
    if not args.no_fp64:
        eng64, fx = run_engine(prob, 0, args.warmup, args.steps, not args.no_autotune, model)
        rl64 = roofline(eng64, prob, args.roofline_reps, PEAK_FP64_TFLOPS)
        eng64.close()
        n = min(len(fx["outs"]), len(mx["outs"]))
        dc = max(abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im))
                 for a, b in zip(mx["outs"][:n], fx["outs"][:n]))
        v64 = fx["fs"] / fx["wall"]
        log(f"max|dC(t)| mixed vs FP64 over {n} output times: {dc:.3e}; fp64 {v64:.4f} fs/s")
        result["accuracy"] = {"max_abs_dC_vs_fp64": dc, "output_times": n,
                              "span_fs": (n - 1) * prob.writetime * AU2FS, "gate": 1e-6, "pass": dc <= 1e-6}
        result["fp64"] = {"value": v64, "unit": "fs/s", "speedup_mixed": value / v64,
                          "HA_ms": rl64["HA_ms"], "roofline_frac": rl64["frac"],
                          "achieved_TFLOPs": rl64["achieved"]}
    if model is not None and not args.no_classical:
        engc, cx = run_engine(prob, 1, args.warmup, args.steps, not args.no_autotune, None, "mixed-classical")
        engc.close()
        result["mixed_classical_controller"] = {
            "value": cx["fs"] / cx["wall"], "unit": "fs/s", "attempts": len(cx["atts"]),
            "rejected": sum(1 for a in cx["atts"] if not a.accepted),
            "max_abs_dC_vs_learned": max(abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im))
                                         for a, b in zip(cx["outs"], mx["outs"]))}
    if not args.no_cpu:
        acc = sorted(a.h for a in mx["atts"] if a.accepted)
        h_med = acc[len(acc) // 2]
        log(f"cpu baseline: one oracle CMF2 attempt at h = {h_med * AU2FS:.4f} fs on {args.cpu_threads} threads")
        result["cpu_baseline"] = cpu_baseline(prob, len(mx["atts"]) / mx["fs"], h_med, args.cpu_threads)
        if args.workload == "c2" and args.c1_fs > 0:
            log(f"cpu baseline: C1 end to end over {args.c1_fs} fs, oracle on {args.cpu_threads} threads and engine")
            result["cpu_baseline_c1"] = c1_baseline(args.c1_fs, args.cpu_threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
