"""Standalone CMF2 run of a Macaw input file on the HIP engine.

    python -m macaw_amd.run input.inp [--precision 0|1] [--controller classical|learned] [--outdir DIR]

Reads the input and operator files (the keyword subset the engine runs),
propagates from tinit to tfinal with the output cadence of Macaw's
propagate_wf (propagator.f90:257-306) and writes, in Macaw's formats:

* ``auto``: header of propagator.f90:26, then per output time
  ``(2x,f10.4,5x,3(1x,f18.14))`` time [fs], Re C, Im C, |C| (output.f90:228-229);
* ``spop`` (more than one state): header of propagator.f90:31, then
  ``(f10.4,2x,30(2x,f10.7))`` time [fs] and the state populations
  (output.f90:243-246);
* ``restart``: the state at the last output time (output.f90:230-241).

The reference propagates one output interval past its last output; this
driver stops at the last output time, whose state the restart file holds.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

AU2FS = 0.02418884326505

AUTO_HEADER = "#    time[fs]         Re(autocorrel)     Im(autocorrel)     Abs(autocorrel)"
SPOP_HEADER = "#  time[fs]      State populations"


def auto_line(t_au: float, re: float, im: float) -> str:
    return f"  {t_au * AU2FS:10.4f}     " + "".join(f" {v:18.14f}" for v in (re, im, float(np.hypot(re, im))))


def spop_line(t_au: float, pops) -> str:
    return f"{t_au * AU2FS:10.4f}  " + "".join(f"  {p:10.7f}" for p in list(pops)[:30])


def write_auto(path: str, outs) -> None:
    with open(path, "w") as fh:
        fh.write(AUTO_HEADER + "\n")
        for o in outs:
            fh.write(auto_line(o.time, o.auto_re, o.auto_im) + "\n")


def write_spop(path: str, outs, pops) -> None:
    with open(path, "w") as fh:
        fh.write(SPOP_HEADER + "\n")
        for o, p in zip(outs, pops):
            fh.write(spop_line(o.time, p) + "\n")


def run(input_path: str, outdir: str = ".", precision: int = 0, controller: str = "classical", log=print):
    from .engine import Engine
    from .problem import load_problem
    from .restart import write_restart
    prob = load_problem(input_path)
    e = Engine(prob, precision=precision)
    if controller == "learned":
        e.set_step_model(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "step_controller.npy")))
    wt = prob.writetime
    t0, tf = prob.tinit, prob.tfinal
    outs, pops, atts = [], np.zeros((0, prob.nstates)), []
    t_end = t0
    if tf - t0 >= wt * 0.9:
        # outputs at t0 .. tf - wt; the state then stands at tf
        outs, pops, atts = e.propagate(t0, tf - wt)
        t_end = outs[-1].time + wt
    last, lp = e.observables(t_end)
    outs = list(outs) + [last]
    pops = np.vstack([pops, lp[None, :]])
    os.makedirs(outdir, exist_ok=True)
    write_auto(os.path.join(outdir, "auto"), outs)
    if prob.nstates > 1:
        write_spop(os.path.join(outdir, "spop"), outs, pops)
    A, psi = e.get_state()
    write_restart(os.path.join(outdir, "restart"), prob, t_end, A, psi)
    e.close()
    rej = sum(1 for a in atts if not a.accepted)
    log(f"[run] {len(outs)} outputs to {t_end * AU2FS:.4f} fs, {len(atts)} CMF2 attempts ({rej} rejected)")
    return outs, pops


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("input")
    ap.add_argument("--outdir", default=".")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1))
    ap.add_argument("--controller", default="classical", choices=("classical", "learned"))
    a = ap.parse_args(argv)
    run(a.input, a.outdir, a.precision, a.controller, log=lambda m: print(m, flush=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
