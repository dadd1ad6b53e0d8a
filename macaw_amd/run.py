"""Standalone CMF2 run of a Macaw input file on the HIP engine.

    python -m macaw_amd.run input.inp [--precision 0|1] [--controller classical|learned] [--outdir DIR]

Reads the input and operator files (the keyword subset the engine runs),
propagates from tinit to tfinal with the output cadence of Macaw's
propagate_wf (CMF2, propagator.f90:257-306, or the variable-mean-field
Runge-Kutta branch, :108-140, for ``integration = rk45 / dopri``) and writes,
in Macaw's formats:

* ``auto``: header of propagator.f90:26, then per output time
  ``(2x,f10.4,5x,3(1x,f18.14))`` time [fs], Re C, Im C, |C| (output.f90:228-229);
* ``spop`` (more than one state): header of propagator.f90:31, then
  ``(f10.4,2x,30(2x,f10.7))`` time [fs] and the state populations
  (output.f90:243-246);
* ``restart``: the state at the last output time (output.f90:230-241);
* ``psi`` (``--psi-every N``): the state (time, A, psi) as one unformatted
  record every N output times from the first (propagator.f90:40-45,
  294-298, output.f90:248-252);
* ``steps`` (``--steps``): header of propagator.f90:37-38, then per CMF2 step
  ``(f12.6,3(2x,f12.6))`` time, attempted, done and suggested step [fs]
  (propagator.f90:273-274);
* ``expectation`` (``expect = file`` in the input): per output time
  ``(1x,f8.2,2x,800(3x,f14.9,1x,f14.9))`` time [fs] and <Psi|O|Psi> of every
  expectation operator (output.f90:449-498);
* ``output``: the propagation interval and the ``psi`` period
  (propagator.f90:71-84), then per output time the block of
  output.f90:152-224 (time, norm, energy and its change from the start,
  electronic populations, natural populations of every mode, and for
  Gaussian modes the gross Gaussian populations and the population errors of
  dropping one Gaussian), followed with ``density`` by the populations of
  the DVR basis functions (output.f90:277-307).  The echo of the input that
  precedes them in Macaw's file is not written;
* ``densmat`` (``densmat = m``), ``density`` (``density``) and ``wpOvl``
  (``wp_ovl``): per output time, as macaw_amd/densities.py describes;
* ``log``: the first tables of Macaw's log (input.f90:391-413).

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


def _mode_lines(label_values):
    """'(a,i3,a,6(2x,f8.5))' and '(a,6(2x,f8.5))' continuation lines."""
    lines = []
    for im, vals in label_values:
        vals = list(vals)
        lines.append(f"    Mode {im:3d}:" + "".join(f"  {v:8.5f}" for v in vals[:6]))
        for j in range(6, len(vals), 6):
            lines.append("             " + "".join(f"  {v:8.5f}" for v in vals[j:j + 6]))
    return lines


def mode_quantities(prob, rho, psi):
    """Natural populations of every (state, mode), and for Gaussian modes the
    gross Gaussian populations and population errors (output.f90:94-135).
    `rho[s, m]` as the engine returns it (the transpose of rho(j1, j2))."""
    from .restart import gaussian_overlap
    ns, nm = prob.nstates, prob.nmodes
    ip, _ = prob.ipsiSPF()
    npar = prob.npar()
    nat, ggp, perr = {}, {}, {}
    for s in range(ns):
        for m in range(nm):
            n = int(prob.nSPF[m, s])
            r = rho[s, m][:n, :n].T                     # rho(j1, j2)
            if not prob.gwp_mode(m):
                nat[s, m] = np.sort(np.linalg.eigvalsh(r))[::-1]
                continue
            jel = s if prob.MultiSet[m] else 0
            nj = int(prob.nSPF[m, jel])
            start = int(ip[m, jel]) - 1
            xi = psi[start:start + nj * int(npar[m])].reshape(int(npar[m]), nj, order="F")
            S = gaussian_overlap(prob, m, xi)[:n, :n]   # S00(j1, j2) = <g_j1|g_j2>
            w, V = np.linalg.eigh(S)
            s12 = np.conj((V * np.sqrt(np.maximum(w, 0.0))) @ V.conj().T)   # conjugated square root
            nat[s, m] = np.sort(np.linalg.eigvalsh(s12 @ r @ s12))[::-1]
            g = np.array([sum((r[j2, j1] * S[j2, j1]).real for j2 in range(j1))
                          + sum((r[j1, j2] * S[j1, j2]).real for j2 in range(j1, n)) for j1 in range(n)])
            e = np.abs(2.0 * g - np.diag(r).real)
            ggp[s, m] = g
            perr[s, m] = np.sort(e)[::-1]
    return nat, ggp, perr


def output_header(t0_au: float, tf_au: float, psi_period_au: float = 0.0):
    """The lines propagate_wf writes to `output` before the first output time
    (propagator.f90:71-84): the propagation interval and, when the `psi` file
    is written, its period."""
    lines = ["", f"Propagation from {t0_au * AU2FS:9.3f} fs to {tf_au * AU2FS:9.3f} fs"]
    if psi_period_au > 0.0:
        lines.append(f"The wavefunction is stored every {psi_period_au * AU2FS:7.3f} fs")
    return lines + [""]


def output_block(prob, out, pops, energy_start, nat, ggp, perr):
    """One output time of Macaw's `output` file (output.f90:152-224)."""
    from .inputfile import AU2EV
    lines = ["", f"Time = {out.time * AU2FS:9.3f} fs",
             f" Norm = {float(np.sum(pops)):12.8f} ,    "
             f" Energy = {out.energy * AU2EV:9.4f} eV,    "
             f" Delta-E = {(out.energy - energy_start) * 1000.0 * AU2EV:9.4f} meV"]
    if prob.nstates > 1:
        lines.append(" Electronic populations:" + "".join(f"  {p:9.6f}" for p in pops[:30]))
    lines.append("  Natural populations:")
    for s in range(prob.nstates):
        lines.append(f"   State {s + 1}")
        lines += _mode_lines((m + 1, nat[s, m]) for m in range(prob.nmodes))
    gm = [m for m in range(prob.nmodes) if prob.gwp_mode(m)]
    if gm:
        for title, table in (("  Gross Gaussian populations:", ggp),
                             ("  Errors on the population by neglecting one Gaussian:", perr)):
            lines.append(title)
            for s in range(prob.nstates):
                lines.append(f"   State {s + 1}")
                lines += _mode_lines((m + 1, table[s, m]) for m in gm)
    return lines


def expectation_line(t_au: float, values) -> str:
    """'(1x,f8.2,2x,800(3x, f14.9,1x,f14.9))' of output.f90:497."""
    return f" {t_au * AU2FS:8.2f}  " + "".join(f"   {v.real:14.9f} {v.imag:14.9f}" for v in values)


class Expectations:
    """<Psi|O|Psi> of the input's expectation operators (output.f90:449-498):
    one engine context per operator, holding the operator as its Hamiltonian
    (same layout and tables).  O's matrix elements come from the current SPFs
    (Calc_h1mode); with Gaussian modes the coefficients are taken to the
    basis orthonormalised by D0 first (B = D^H S A, cmf2_base.f90:264-305),
    where <Psi|O|Psi> = <B|D^H O D|B>."""

    def __init__(self, prob):
        from .engine import Engine
        from .problem import with_operator
        self.prob = prob
        self.engines = [Engine(with_operator(prob, terms), precision=0) for _, terms in prob.expect]
        self.gwp = any(prob.gwp_mode(m) for m in range(prob.nmodes))

    def __call__(self, A, psi):
        vals = []
        for e in self.engines:
            e.set_state(A, psi)
            if self.gwp:
                e.prepare_gwp()
                _, D = e.gwp_overlaps()
                B = e.a_to_b(A, D)
            else:
                e.h1mode()
                B = A
            vals.append(complex(np.vdot(B, e.apply_h(B))))
        return vals

    def close(self):
        for e in self.engines:
            e.close()


STEPS_HEADER = ["#                Integration steps  [fs]", "#  time [fs]      attempted       done       suggested"]


def steps_line(t, attempted, done, suggested) -> str:
    """'(f12.6,3(2x,f12.6))' of propagator.f90:274, in fs."""
    return f"{t * AU2FS:12.6f}" + "".join(f"  {v * AU2FS:12.6f}" for v in (attempted, done, suggested))


def _interval_by_steps(e, prob, t, h, steps):
    """One output interval as propagate_wf's CMF2 loop (propagator.f90:270-293)
    through single CMF2 steps from step size h, recording (time, attempted,
    done, suggested); the arithmetic is the engine's own interval loop, step
    for step.  Returns the new time and step size."""
    wt, minstep = prob.writetime, prob.min_stepsize
    tnext = t + wt
    while True:
        done, nxt = e.cmf2_step(t, h)
        steps.append((t, h, done, nxt))
        t = t + done
        if t >= tnext - minstep:
            break
        k = int((tnext - t) / nxt)
        h = (tnext - t) if k == 0 else (tnext - t) / (k + 1)
    return t, h


def run(input_path: str, outdir: str = ".", precision: int = 0, controller: str = "classical", log=print,
        steps: bool = False, psi_every: int = 0, autotune: bool = False):
    from .engine import Engine
    from .problem import load_problem
    from . import densities as DENS
    from .restart import _write_record as write_record, write_psi_record, write_restart
    prob = load_problem(input_path)
    if steps and prob.normalize:
        raise ValueError("--steps drives single CMF2 steps from the host and does not renormalise; "
                         "run without --steps when the input sets `normalize`")
    e = Engine(prob, precision=precision)
    e.set_normalize(prob.normalize)
    if controller == "learned":
        e.set_step_model(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "step_controller.npy")))
    if autotune:
        brick, ms = e.autotune()
        log(f"[run] autotune: LDS brick {brick} ({ms:.2f} ms per H.A + mean fields)")
    wt = prob.writetime
    t0, tf = prob.tinit, prob.tfinal
    # integration = rk45 / dopri: variable mean field (propagator.f90:108-140)
    vmf = prob.int_type in (1, 2)
    os.makedirs(outdir, exist_ok=True)
    outs, pops, atts, step_log = [], [], [], []
    block = output_header(t0, tf, psi_every * wt)
    energy_start = None
    t = t0
    h = e.stepsize()                            # writetime * eps_int (propagator.f90:103)
    # one output interval per call (the step size carries over between
    # calls), so that rho of every output time is at hand for `output`
    psi_fh = open(os.path.join(outdir, "psi"), "wb") if psi_every > 0 else None
    ex = prob.extra
    with open(os.path.join(outdir, "log"), "w") as fh:
        fh.write("\n".join(DENS.log_tables(prob, ex["grids"], ex["hamiltonian_file"])) + "\n")
    dm_fh = open(os.path.join(outdir, "densmat"), "wb") if ex.get("densmat") else None
    dens_fh = open(os.path.join(outdir, "density"), "wb") if ex.get("density") else None
    wpovl = [] if ex.get("wp_ovl") else None
    expect = Expectations(prob) if prob.expect else None
    expect_lines = []
    k_out = 0
    while True:
        rho = e.rho()
        A_t, psi = e.get_state()
        if expect is not None:
            expect_lines.append(expectation_line(t, expect(A_t, psi)))
        if psi_fh is not None and k_out % psi_every == 0:
            write_psi_record(psi_fh, t, A_t, psi)
        k_out += 1
        last = t >= tf - wt * 0.1
        if last or (steps and not vmf):
            o, p = e.observables(t)
            a = []
        elif vmf:
            ol, pl, a = e.propagate_vmf(prob.int_type, t, t)
            o, p = ol[0], pl[0]
        else:
            ol, pl, a = e.propagate(t, t)
            o, p = ol[0], pl[0]
        if energy_start is None:
            energy_start = o.energy * o.norm      # energyMF at the start (propagator.f90:96)
        block += output_block(prob, o, p, energy_start, *mode_quantities(prob, rho, psi))
        if dm_fh is not None:
            DENS.write_densmat(dm_fh, prob, t, rho, psi, ex["densmat"])
        if dens_fh is not None:
            recs, lines = DENS.density_records(prob, t, rho, psi, ex["trafo"])
            for rec in recs:
                write_record(dens_fh, rec)
            block += lines
        if wpovl is not None:
            # one wave packet: the 1 x 1 overlap matrix <Psi|Psi> (Calc_SWP)
            wpovl += DENS.wpovl_lines(t, complex(float(np.sum(p)), 0.0), header=not wpovl)
        outs.append(o)
        pops.append(p)
        atts += list(a)
        if last:
            break
        if steps and not vmf:
            t, h = _interval_by_steps(e, prob, t, h, step_log)
        else:
            acc = [x for x in a if x.accepted]
            t = acc[-1].time + acc[-1].h        # the engine's own accumulated time
    t_end = t
    for fh in (psi_fh, dm_fh, dens_fh):
        if fh is not None:
            fh.close()
    if wpovl is not None:
        with open(os.path.join(outdir, "wpOvl"), "w") as fh:
            fh.write("\n".join(wpovl) + "\n")
    pops = np.asarray(pops)
    with open(os.path.join(outdir, "output"), "w") as fh:
        fh.write("\n".join(block) + "\n")
    write_auto(os.path.join(outdir, "auto"), outs)
    if expect is not None:
        expect.close()
        with open(os.path.join(outdir, "expectation"), "w") as fh:
            fh.write("\n".join(expect_lines) + "\n")
    if prob.nstates > 1:
        write_spop(os.path.join(outdir, "spop"), outs, pops)
    if steps:
        with open(os.path.join(outdir, "steps"), "w") as fh:
            fh.write("\n".join(STEPS_HEADER + [steps_line(*r) for r in step_log]) + "\n")
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
    ap.add_argument("--steps", action="store_true", help="also write Macaw's `steps` file (one line per CMF2 step)")
    ap.add_argument("--autotune", action="store_true", help="choose the LDS brick size by timing the candidates")
    ap.add_argument("--psi-every", type=int, default=0,
                    help="write the state to Macaw's `psi` file every N output times (0: never)")
    a = ap.parse_args(argv)
    run(a.input, a.outdir, a.precision, a.controller, log=lambda m: print(m, flush=True), steps=a.steps,
        psi_every=a.psi_every, autotune=a.autotune)
    return 0


if __name__ == "__main__":
    sys.exit(main())
