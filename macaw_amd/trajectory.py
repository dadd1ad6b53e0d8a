"""Long propagations of a workload with their observables and attempt logs on
disk, and the precision comparison of such runs.

    python -m macaw_amd.trajectory run --workload c2 --precision 0 --out DIR [--t-end FS] [--budget S]
    python -m macaw_amd.trajectory run --workload c2 --precision 0 --out DIR --resume [--budget S]
    python -m macaw_amd.trajectory run --workload c2 --precision 1 --replay DIR64/attempts.txt --out DIR
    python -m macaw_amd.trajectory compare --ref DIR64 --runs DIR_A DIR_B ... [--md FILE]

`run` propagates from t = 0 one output interval per engine call (the step
size and the controller state carry over, so the trajectory is the single
call's) and appends after every interval: the output record (time, C(t),
norm, energy, state populations) to DIR/outputs.txt and the interval's CMF2
attempts to DIR/attempts.txt, each line flushed, so an interrupted run keeps
every finished interval.  It stops at --t-end or when the next interval would
pass the wall budget (the mean interval time so far).  With --replay the
engine takes the step sizes and accept/reject decisions of that attempt log
(mg_set_replay), which removes the controller from a precision comparison;
the run stops where the log ends.  When a run stops on its budget it saves
its state (time, A, SPFs, the step size reached) to DIR/state.npz; --resume
continues from there, appending to the same files, so a long run can be
split over several processes.  A resumed segment starts the learned
controller's step history afresh (the step size itself carries over).

`compare` reads runs and reports, against the reference run, the largest
differences in C(t), populations, norm and energy over the common output
times, their growth with time and each run's propagated fs per wall second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

AU2FS = 0.02418884326505
ATT_FIELDS = ("time", "h", "err_A", "err_psi", "accepted", "nkrylov1", "nkrylov2")


def _log(msg):
    print(f"[trajectory] {msg}", flush=True)


def build(workload):
    from .lvc import make_problem
    with tempfile.TemporaryDirectory() as d:
        return make_problem(workload, d)


def read_attempts(path):
    from .abi import MgAttempt
    out = []
    with open(path) as fh:
        for line in fh:
            if not line.strip() or line.startswith("#"):
                continue
            v = line.split()
            out.append(MgAttempt(time=float(v[0]), h=float(v[1]), err_A=float(v[2]), err_psi=float(v[3]),
                                 accepted=int(v[4]), nkrylov1=int(v[5]), nkrylov2=int(v[6])))
    return out


def read_outputs(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.strip() and not line.startswith("#"):
                rows.append([float(x) for x in line.split()])
    return np.array(rows)


def run(args):
    from .engine import Engine
    prob = build(args.workload)
    os.makedirs(args.out, exist_ok=True)
    eng = Engine(prob, precision=args.precision)
    state_path = args.state or os.path.join(args.out, "state.npz")
    resume = None
    if args.resume:
        if args.replay:
            raise SystemExit("--resume and --replay cannot be combined")
        with np.load(state_path) as z:
            resume = {k: z[k] for k in z.files}
        with open(os.path.join(args.out, "meta.json")) as fh:
            prev = json.load(fh)
        if (prev["workload"], prev["precision"], prev["controller"]) != (args.workload, args.precision,
                                                                        args.controller):
            raise SystemExit(f"--resume: {args.out} holds a run of another workload, precision or controller")
    if args.controller == "learned":
        eng.set_step_model(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "step_controller.npy")))
    t_log_end = None
    if args.replay:
        log = read_attempts(args.replay)
        eng.set_replay(log)
        acc = [a for a in log if a.accepted]
        t_log_end = acc[-1].time + acc[-1].h
        _log(f"replaying {len(log)} attempts to {t_log_end * AU2FS:.3f} fs")
    if args.autotune:
        brick, ms = eng.autotune()
        _log(f"autotune: brick {brick} ({ms:.2f} ms)")
    tout = prob.writetime
    t_end = args.t_end / AU2FS
    if t_log_end is not None:
        t_end = min(t_end, t_log_end)
    meta = {"workload": args.workload, "precision": args.precision, "controller": args.controller,
            "replay": args.replay or None, "budget_s": args.budget, "t_end_fs": args.t_end}
    t, k, natt, wall_prev, segments = 0.0, 0, 0, 0.0, 1
    if resume is not None:
        meta = prev
        meta["t_end_fs"] = args.t_end
        t = float(resume["t"])
        eng.set_state(resume["A"], resume["psi"])
        eng.set_stepsize(float(resume["stepsize"]))
        natt, wall_prev = int(prev.get("attempts", 0)), float(prev.get("wall_s", 0.0))
        segments = int(prev.get("segments", 1)) + 1
        _log(f"resuming at {t * AU2FS:.3f} fs (segment {segments}), step {float(resume['stepsize']) * AU2FS:.3e} fs")
    with open(os.path.join(args.out, "meta.json"), "w") as fh:
        json.dump(meta, fh)
    mode = "a" if resume is not None else "w"
    fo = open(os.path.join(args.out, "outputs.txt"), mode)
    fa = open(os.path.join(args.out, "attempts.txt"), mode)
    if resume is None:
        fo.write("# time_fs  Re_C  Im_C  norm  energy_au  " + "  ".join(f"pop{s + 1}" for s in range(prob.nstates))
                 + "\n")
        fa.write("# " + "  ".join(ATT_FIELDS) + "\n")
    wall0 = time.perf_counter()
    # the record at the resume time was written when the previous segment stopped
    skip_record = resume is not None
    while True:
        last = t >= t_end - 0.1 * tout
        t_int = time.perf_counter()
        if last:
            o, p = eng.observables(t)
            atts = []
        else:
            ol, pl, atts = eng.propagate(t, t)
            o, p = ol[0], pl[0]
        if not skip_record:
            fo.write(f"{o.time * AU2FS:.10f} {o.auto_re:.17e} {o.auto_im:.17e} {o.norm:.17e} {o.energy:.17e} "
                     + " ".join(f"{x:.17e}" for x in p) + "\n")
            fo.flush()
        skip_record = False
        for a in atts:
            fa.write(f"{a.time!r} {a.h!r} {a.err_A!r} {a.err_psi!r} {a.accepted} {a.nkrylov1} {a.nkrylov2}\n")
        fa.flush()
        natt += len(atts)
        if last:
            break
        acc = [a for a in atts if a.accepted]
        t = acc[-1].time + acc[-1].h
        k += 1
        el = time.perf_counter() - wall0
        _log(f"{t * AU2FS:.2f} fs  {el:.0f} s  {natt} attempts  interval {time.perf_counter() - t_int:.1f} s  "
             f"C = {o.auto_re:+.8f}{o.auto_im:+.8f}i")
        if el + el / k > args.budget or (args.max_intervals and k >= args.max_intervals):
            # one more output record at the time reached, and the state to
            # resume from, then stop
            o, p = eng.observables(t)
            fo.write(f"{o.time * AU2FS:.10f} {o.auto_re:.17e} {o.auto_im:.17e} {o.norm:.17e} {o.energy:.17e} "
                     + " ".join(f"{x:.17e}" for x in p) + "\n")
            fo.flush()
            if not args.replay and state_path != "none":
                A, psi = eng.get_state()
                tmp = state_path + ".tmp.npz"
                np.savez(tmp, t=t, A=A, psi=psi, stepsize=eng.stepsize())
                os.replace(tmp, state_path)
            break
    wall = wall_prev + time.perf_counter() - wall0
    meta.update({"t_reached_fs": t * AU2FS, "wall_s": wall, "attempts": natt, "fs_per_s": t * AU2FS / wall,
                 "segments": segments})
    with open(os.path.join(args.out, "meta.json"), "w") as fh:
        json.dump(meta, fh)
    fo.close()
    fa.close()
    eng.close()
    _log(f"done: {t * AU2FS:.2f} fs in {wall:.0f} s ({t * AU2FS / wall:.4f} fs/s), {natt} attempts")


def compare(args):
    ref = read_outputs(os.path.join(args.ref, "outputs.txt"))
    mref = json.load(open(os.path.join(args.ref, "meta.json")))
    lines = [f"reference: {args.ref} (precision {mref['precision']}, {mref.get('t_reached_fs', 0):.2f} fs, "
             f"{mref.get('fs_per_s', 0):.4f} fs/s, {mref.get('attempts', 0)} attempts)", ""]
    lines.append("| run | precision | steps | span (fs) | fs/s | attempts | max abs dC | max abs dpop | "
                 "max abs dnorm | max abs dE (eV) | max abs dC, first half | max abs dC, second half |")
    lines.append("|---|---|---|---|---|---|---|---|---|---|---|---|")
    curves = {}
    for d in args.runs:
        out = read_outputs(os.path.join(d, "outputs.txt"))
        meta = json.load(open(os.path.join(d, "meta.json")))
        # common output times (the records at multiples of the interval)
        n = 0
        while n < min(len(out), len(ref)) and abs(out[n, 0] - ref[n, 0]) < 1e-6:
            n += 1
        c_r = ref[:n, 1] + 1j * ref[:n, 2]
        c_o = out[:n, 1] + 1j * out[:n, 2]
        dc = np.abs(c_o - c_r)
        dpop = np.abs(out[:n, 5:] - ref[:n, 5:]).max(axis=1)
        dnorm = np.abs(out[:n, 3] - ref[:n, 3])
        de = np.abs(out[:n, 4] - ref[:n, 4]) * 27.2113961
        h = max(1, n // 2)
        curves[d] = (ref[:n, 0], dc)
        lines.append(f"| {d} | {meta['precision']} | {'replayed' if meta.get('replay') else 'free'} | "
                     f"{ref[n - 1, 0]:.1f} | {meta.get('fs_per_s', 0):.4f} | {meta.get('attempts', 0)} | "
                     f"{dc.max():.2e} | {dpop.max():.2e} | {dnorm.max():.2e} | {de.max():.2e} | "
                     f"{dc[:h].max():.2e} | {dc[h:].max() if n > h else 0.0:.2e} |")
    lines.append("")
    lines.append("max abs dC(t) at selected times (fs):")
    for d, (tt, dc) in curves.items():
        marks = [i for i in range(len(tt)) if abs(tt[i] - round(tt[i] / 10.0) * 10.0) < 1e-6]
        lines.append(f"- {d}: " + ", ".join(f"{tt[i]:.0f}: {np.abs(dc[:i + 1]).max():.2e}" for i in marks))
    text = "\n".join(lines)
    print(text)
    if args.md:
        with open(args.md, "w") as fh:
            fh.write(text + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", default="c2")
    r.add_argument("--precision", type=int, default=0, choices=(0, 1))
    r.add_argument("--controller", default="learned", choices=("learned", "classical"))
    r.add_argument("--t-end", type=float, default=200.0, help="fs")
    r.add_argument("--budget", type=float, default=1e9, help="wall seconds")
    r.add_argument("--replay", help="attempts.txt of another run to replay")
    r.add_argument("--autotune", action="store_true")
    r.add_argument("--resume", action="store_true", help="continue the run saved in --out")
    r.add_argument("--state", help="file of the saved state (default OUT/state.npz; 'none': not saved)")
    r.add_argument("--max-intervals", type=int, default=0,
                   help="stop (and save the state) after this many output intervals of this segment")
    r.add_argument("--out", required=True)
    c = sub.add_parser("compare")
    c.add_argument("--ref", required=True)
    c.add_argument("--runs", nargs="+", required=True)
    c.add_argument("--md")
    a = ap.parse_args(argv)
    if a.cmd == "run":
        run(a)
    else:
        compare(a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
