"""Learned CMF2 step-size controller: training data from the engine's own
step logs, training on the device, evaluation against the classical rule.

The classical controller (Macaw's cmf2, cmf2_base.f90:178-199) proposes
h_next = min(h / sqrt(err), 1.5 h) after an accepted step; about a quarter of
the attempts it starts on the LVC workloads are then rejected and repeated.
The learned controller replaces only that proposal: after each accepted step
the engine evaluates a 12-64-64-1 perceptron on 12 step features (see
``mg_step_features`` in include/macaw_gpu.h) and takes h exp(y).  The CMF2
error estimate still accepts or rejects every attempt, so the accuracy
criterion is unchanged.

Training pairs come from plain (classical) runs: the features after accepted
step n and the first attempt of step n+1, (h', err').  With the local error
model err ~ h^p (p estimated from the rejected-then-retried attempts of the
same logs, median, clipped to [1, 4]), the step that would have produced the
target error ratio r* = 0.81 is h* = h' (r* / err')^(1/p); the regression
target is y = log(h* / h_n), clipped to the controller's range
[log 0.2, log 2].

Command line:
  python -m macaw_amd.controller collect --spec c1 --seeds 8 --fs 20 [--precision 1] --out data.npz
  python -m macaw_amd.controller train --data data.npz --out model.npy
  python -m macaw_amd.controller evaluate --model model.npy --spec c1 --seed 999 --fs 20
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

AU2FS = 0.02418884326505
Y_MIN, Y_MAX = math.log(0.2), math.log(2.0)


def error_exponent(attempts, eps_int, default=2.0):
    """Median exponent p of err ~ h^p over (rejected, retried) attempt pairs."""
    est = []
    for a, b in zip(attempts[:-1], attempts[1:]):
        if a.accepted or b.h >= a.h:
            continue
        ea, eb = max(a.err_A, a.err_psi) / eps_int, max(b.err_A, b.err_psi) / eps_int
        if ea > 0 and eb > 0:
            est.append(math.log(ea / eb) / math.log(a.h / b.h))
    if not est:
        return default
    return float(min(4.0, max(1.0, np.median(est))))


def training_pairs(features, h_acc, attempts, eps_int, exponent=None, target=0.81):
    """(X, y) pairs from one logged propagation: feature row n belongs to the
    n-th accepted attempt; its target comes from the attempt after it."""
    p = error_exponent(attempts, eps_int) if exponent is None else exponent
    acc = [i for i, a in enumerate(attempts) if a.accepted]
    if len(acc) != len(features):
        raise ValueError(f"{len(features)} feature rows for {len(acc)} accepted attempts")
    X, y = [], []
    for row, i in enumerate(acc):
        if i + 1 >= len(attempts):
            break
        nxt = attempts[i + 1]
        e = max(nxt.err_A, nxt.err_psi) / eps_int
        if not (e > 0):
            continue
        h_star = nxt.h * (target / e) ** (1.0 / p)
        y.append(min(Y_MAX, max(Y_MIN, math.log(h_star / h_acc[row]))))
        X.append(features[row])
    return np.asarray(X, np.float32).reshape(-1, 12), np.asarray(y, np.float32), p


def seeded_spec(name, seed):
    from .lvc import SPECS
    return dataclasses.replace(SPECS[name], seed=seed, name=f"{name}_s{seed}")


def collect(spec_name, seeds, t_fs, precision=0, log=print):
    """Classical runs of `spec_name` with fresh seeds; returns X, y."""
    from .engine import Engine
    from .lvc import make_problem
    Xs, ys = [], []
    for seed in seeds:
        with tempfile.TemporaryDirectory() as d:
            prob = make_problem(seeded_spec(spec_name, seed), d)
        e = Engine(prob, precision=precision)
        e.log_step_features(True)
        t0 = time.perf_counter()
        _, _, atts = e.propagate(0.0, t_fs / AU2FS)
        f, h = e.step_features()
        X, y, p = training_pairs(f, h, atts, prob.eps_int)
        rej = sum(1 for a in atts if not a.accepted)
        log(f"[collect] {spec_name} (precision {precision}) seed {seed}: {len(atts)} attempts ({rej} rejected), "
            f"{len(y)} pairs, "
            f"p = {p:.2f}, {time.perf_counter() - t0:.1f} s")
        Xs.append(X)
        ys.append(y)
        e.close()
    return np.concatenate(Xs), np.concatenate(ys)


def run(prob, params, t_fs, precision=0):
    from .engine import Engine
    e = Engine(prob, precision=precision)
    if params is not None:
        e.set_step_model(params)
    t0 = time.perf_counter()
    outs, pops, atts = e.propagate(0.0, t_fs / AU2FS)
    wall = time.perf_counter() - t0
    stats = e.controller_stats() if params is not None else (0, 0)
    e.close()
    return outs, pops, atts, wall, stats


def evaluate(prob, params, t_fs, ref_factor=0.5, precision=0, log=print):
    """Classical vs learned controller on one problem (engine precision
    `precision`), both against an FP64 reference run at eps_int * ref_factor
    (the DeltaPsi estimate has a rounding floor near 2e-8, so much tighter
    references stall)."""
    ref_prob = dataclasses.replace(prob, eps_int=prob.eps_int * ref_factor)
    ref, _, _, _, _ = run(ref_prob, None, t_fs)
    res = {}
    for tag, prm in (("classical", None), ("learned", params)):
        outs, pops, atts, wall, stats = run(prob, prm, t_fs, precision)
        dc = max(abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im)) for a, b in zip(outs, ref))
        rej = sum(1 for a in atts if not a.accepted)
        res[tag] = {"attempts": len(atts), "rejected": rej, "rejected_fraction": rej / max(1, len(atts)),
                    "wall_s": wall, "max_dC_vs_reference": dc, "reference_eps_int": ref_prob.eps_int,
                    "model_proposals": stats[0],
                    "model_fallbacks": stats[1]}
        log(f"[evaluate] {tag}: {len(atts)} attempts, {rej} rejected ({rej / max(1, len(atts)):.1%}), "
            f"{wall:.2f} s, max|dC| vs the eps_int x {ref_factor} run {dc:.2e}")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--spec", default="c1")
    c.add_argument("--seeds", type=int, default=8)
    c.add_argument("--seed0", type=int, default=1000)
    c.add_argument("--fs", type=float, default=20.0)
    c.add_argument("--precision", type=int, default=0, help="0 complex FP64, 1 mixed")
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--repeat", type=int, nargs="*", default=[],
                   help="how many times each --data file enters the training set (default 1 each)")
    v = sub.add_parser("evaluate")
    v.add_argument("--model", required=True)
    v.add_argument("--spec", default="c1")
    v.add_argument("--seed", type=int, default=999)
    v.add_argument("--fs", type=float, default=20.0)
    v.add_argument("--ref-factor", type=float, default=0.5)
    v.add_argument("--precision", type=int, default=0, help="0 complex FP64, 1 mixed")
    v.add_argument("--json")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "collect":
        X, y = collect(a.spec, range(a.seed0, a.seed0 + a.seeds), a.fs, precision=a.precision, log=log)
        np.savez(a.out, X=X, y=y)
        log(f"[collect] {len(y)} pairs -> {a.out}")
    elif a.cmd == "train":
        from .engine import Engine
        from .lvc import make_problem
        parts = [np.load(p) for p in a.data]
        rep = list(a.repeat) + [1] * (len(parts) - len(a.repeat))
        X = np.concatenate([q["X"] for q, r in zip(parts, rep) for _ in range(r)])
        y = np.concatenate([q["y"] for q, r in zip(parts, rep) for _ in range(r)])
        with tempfile.TemporaryDirectory() as d:
            e = Engine(make_problem("t1", d))
        t0 = time.perf_counter()
        params, loss = e.mlp_train(X, y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
        log(f"[train] {len(y)} pairs, {a.epochs} epochs: mse {loss:.4f} (target variance {np.var(y):.4f}) "
            f"in {time.perf_counter() - t0:.1f} s")
        np.save(a.out, params)
    else:
        from .lvc import make_problem
        params = np.load(a.model)
        with tempfile.TemporaryDirectory() as d:
            prob = make_problem(seeded_spec(a.spec, a.seed), d)
        res = evaluate(prob, params, a.fs, ref_factor=a.ref_factor, precision=a.precision, log=log)
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(res, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
