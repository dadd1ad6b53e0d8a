"""Macaw's output formats written by the standalone driver (macaw_amd/run.py):
`auto` and `spop` lines against the same values formatted by a Fortran
program (ROCm flang) with the reference's edit descriptors (output.f90:228,
244); on the GPU, the standalone run against the Fortran driver's files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from macaw_amd import run as RUN

FLANG = "/opt/rocm/lib/llvm/bin/flang"
AU2FS = 0.02418884326505

FORTRAN = """
program fmt
  implicit none
  integer :: i, n
  double precision :: t, re, im, p(3)
  read(*,*) n
  do i = 1, n
     read(*,*) t, re, im, p
     write(*,'(2x,f10.4,5x,3(1x,f18.14))') t, re, im, sqrt(re**2 + im**2)
     write(*,'(f10.4,2x,30(2x,f10.7))') t, p
  end do
end program fmt
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_lines_match_fortran_edit_descriptors(tmp_path):
    src = tmp_path / "fmt.f90"
    src.write_text(FORTRAN)
    exe = tmp_path / "fmt"
    subprocess.check_call([FLANG, "-O0", "-o", str(exe), str(src)], cwd=str(tmp_path))
    rng = np.random.default_rng(7)
    rows = [(0.0, 1.0, 0.0, (1.0, 0.0, 0.0)), (12.5, -0.25, 0.5, (0.1, 0.2, 0.7))]
    for _ in range(20):
        t = rng.uniform(0, 500)
        re, im = rng.uniform(-1, 1, size=2)
        p = rng.dirichlet([1, 1, 1])
        rows.append((t, re, im, tuple(p)))
    rows = [(float(t), float(re), float(im), tuple(float(x) for x in p)) for t, re, im, p in rows]
    text = f"{len(rows)}\n" + "".join(f"{t!r} {re!r} {im!r} {p[0]!r} {p[1]!r} {p[2]!r}\n" for t, re, im, p in rows)
    out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.splitlines()
    for k, (t, re, im, p) in enumerate(rows):
        assert RUN.auto_line(t / AU2FS, re, im) == out[2 * k], k
        assert RUN.spop_line(t / AU2FS, p) == out[2 * k + 1], k


FORTRAN_OUTPUT = """
program outfmt
  implicit none
  integer :: j1, j2, im, nS
  double precision :: t, norm, e, de, pop(3), natpop(9)
  read(*,*) t, norm, e, de, pop, natpop
  write(*,'(a,f9.3,a)') 'Time = ', t, ' fs'
  write(*,'(a,f12.8,a,4x,a,f9.4,a,4x,a,f9.4,a)') ' Norm = ', norm, ' ,', ' Energy = ', e, ' eV,', ' Delta-E = ', de, ' meV'
  write(*,'(a,30(2x,f9.6))') ' Electronic populations:', pop
  im = 2
  nS = 9
  write(*,'(a,i3,a,6(2x,f8.5))') '    Mode ', im, ':', (natpop(j1), j1 = 1, min(nS,6))
  j2 = 7
  do
     if (j2 .gt. nS) exit
     write(*,'(a,6(2x,f8.5))') '             ', (natpop(j1), j1 = j2, min(nS,j2 + 5))
     j2 = j2 + 6
  end do
end program outfmt
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_output_block_matches_fortran_edit_descriptors(tmp_path):
    import types
    src = tmp_path / "outfmt.f90"
    src.write_text(FORTRAN_OUTPUT)
    exe = tmp_path / "outfmt"
    subprocess.check_call([FLANG, "-O0", "-o", str(exe), str(src)], cwd=str(tmp_path))
    AU2EV = 27.2113961
    t, norm_pops = 37.25, (0.25, 0.5, 0.2499999)
    energy, energy_start = 0.031415, 0.0314
    nat = [0.9, 0.05, 0.03, 0.01, 0.005, 0.003, 0.001, 0.0005, 0.0005]
    vals = [t, sum(norm_pops), energy * AU2EV, (energy - energy_start) * 1000 * AU2EV, *norm_pops, *nat]
    out = subprocess.run([str(exe)], input=" ".join(repr(float(v)) for v in vals) + "\n", capture_output=True,
                         text=True, check=True).stdout.splitlines()
    prob = types.SimpleNamespace(nstates=3, nmodes=2, gwp_mode=lambda m: False)
    o = types.SimpleNamespace(time=t / AU2FS, energy=energy)
    lines = RUN.output_block(prob, o, np.array(norm_pops), energy_start,
                             {(s, m): nat for s in range(3) for m in range(2)}, {}, {})
    assert lines[0] == ""
    assert lines[1:4] == out[:3]
    mode2 = lines.index("   State 1") + 3             # state 1: mode 1 takes 2 lines (6 + 3 values)
    assert lines[mode2: mode2 + 2] == out[3:5]


FORTRAN_HEADER = """
program hdr
  implicit none
  double precision :: t0, tf, tpsi
  read(*,*) t0, tf, tpsi
  write(*,*)
  write(*,'(a,f9.3,a,f9.3,a)') 'Propagation from ', t0, ' fs to ', tf, ' fs'
  write(*,'(a,f7.3,a)') 'The wavefunction is stored every ', tpsi, ' fs'
  write(*,*)
end program hdr
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_output_header_matches_fortran_edit_descriptors(tmp_path):
    src = tmp_path / "hdr.f90"
    src.write_text(FORTRAN_HEADER)
    exe = tmp_path / "hdr"
    subprocess.check_call([FLANG, "-O0", "-o", str(exe), str(src)], cwd=str(tmp_path))
    t0, tf, tpsi = 0.0, 123.45678, 2.5
    out = subprocess.run([str(exe)], input=f"{t0!r} {tf!r} {tpsi!r}\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")[:-1]
    assert RUN.output_header(t0 / AU2FS, tf / AU2FS, tpsi / AU2FS) == out
    assert RUN.output_header(t0 / AU2FS, tf / AU2FS) == out[:2] + out[3:]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_standalone_run_matches_fortran_driver(tmp_path):
    from test_fortran_driver import build_driver
    from macaw_amd.fortran_io import write_descriptor
    from macaw_amd.lvc import make_problem
    from macaw_amd.restart import read_restart
    exe = build_driver()
    prob = make_problem("t1", str(tmp_path))
    inp = str(tmp_path / "t1.inp")
    outs, pops = RUN.run(inp, str(tmp_path / "py"))
    write_descriptor(prob, str(tmp_path / "t1.bin"), prob.tfinal)
    subprocess.check_call([exe, str(tmp_path / "t1.bin")], cwd=str(tmp_path), timeout=300)
    assert (tmp_path / "py" / "auto").read_text() == (tmp_path / "auto").read_text()
    assert (tmp_path / "py" / "spop").read_text() == (tmp_path / "spop").read_text()
    d = read_restart(str(tmp_path / "py" / "restart"))
    assert abs(d["time"] - outs[-1].time) < 1e-9 and d["A"].size == prob.nAConftot
    text = (tmp_path / "py" / "output").read_text()
    assert text.count("Time = ") == len(outs)
    # natural populations of a mode add up to the state population
    blocks = text.split("Time = ")[1:]
    for blk, p in zip(blocks, pops):
        states = blk.split("   State ")[1:]
        for s, st in enumerate(states):
            for line in st.splitlines()[1:]:
                if line.startswith("    Mode"):
                    v = [float(x) for x in line.split(":")[1].split()]
                    assert abs(sum(v) - p[s]) < 1e-4


@pytest.mark.gpu
def test_standalone_run_gaussian_output_sums(tmp_path):
    """G-MCTDH input: per mode, the natural populations and (Gaussian modes)
    the gross Gaussian populations each add up to the state population
    (tr(S^1/2 rho S^1/2) = sum rho o S)."""
    from macaw_amd.lvc import make_problem
    make_problem("g1", str(tmp_path))
    outs, pops = RUN.run(str(tmp_path / "g1.inp"), str(tmp_path / "py"))
    text = (tmp_path / "py" / "output").read_text()
    assert "Gross Gaussian populations" in text
    blocks = text.split("Time = ")[1:]
    assert len(blocks) == len(outs)
    for blk, p in zip(blocks, pops):
        body = blk.split("  Errors on the population")[0]
        for part in body.split("  Natural populations:")[1].split("  Gross Gaussian populations:"):
            for s, st in enumerate(part.split("   State ")[1:]):
                for line in st.splitlines()[1:]:
                    if line.startswith("    Mode"):
                        v = [float(x) for x in line.split(":")[1].split()]
                        assert abs(sum(v) - p[s]) < 1e-4, line


def test_steps_line_format():
    # '(f12.6,3(2x,f12.6))' in fs
    line = RUN.steps_line(1.0 / AU2FS, 0.01 / AU2FS, 0.01 / AU2FS, 0.015 / AU2FS)
    assert line == "    1.000000      0.010000      0.010000      0.015000"


@pytest.mark.gpu
def test_steps_loop_reproduces_the_engine_loop(tmp_path):
    """`--steps` drives the CMF2 loop one step at a time from the host; its
    trajectory is the engine's own interval loop's (auto byte-equal), and the
    steps file holds one line per accepted step, landing on the output times."""
    from macaw_amd.lvc import make_problem
    make_problem("t1", str(tmp_path))
    RUN.run(str(tmp_path / "t1.inp"), str(tmp_path / "a"))
    RUN.run(str(tmp_path / "t1.inp"), str(tmp_path / "b"), steps=True, psi_every=4)
    assert (tmp_path / "a" / "auto").read_text() == (tmp_path / "b" / "auto").read_text()
    rows = np.loadtxt(str(tmp_path / "b" / "steps"), comments="#")
    assert rows.shape[1] == 4 and len(rows) > 10
    # times printed to 1e-6 fs: non-decreasing (the first steps are ~1e-7 fs)
    assert np.all(np.diff(rows[:, 0]) >= 0) and np.all(rows[:, 2] <= rows[:, 1] + 1e-12)
    assert abs(rows[-1, 0] + rows[-1, 2] - 5.0) < 2e-6          # the last step lands on tfinal
    # psi file: the state at outputs 0, 4 and 8 (0, 2 and 4 fs), the first the initial state
    from macaw_amd.problem import load_problem
    from macaw_amd.restart import read_psi_file
    prob = load_problem(str(tmp_path / "t1.inp"))
    recs = read_psi_file(str(tmp_path / "b" / "psi"), prob.nAConftot, prob.dimpsi)
    assert [round(r[0] * AU2FS, 6) for r in recs] == [0.0, 2.0, 4.0]
    assert np.array_equal(recs[0][1], prob.A0) and np.array_equal(recs[0][2], prob.psi0)


@pytest.mark.gpu
@pytest.mark.parametrize("method,name", [(2, "dopri"), (1, "rk45")])
def test_standalone_vmf_input_matches_single_call(method, name, tmp_path):
    """An input with `integration = dopri / rk45` runs the variable-mean-field
    propagator; one output interval per call reproduces a single call."""
    import re
    from macaw_amd.engine import Engine
    from macaw_amd.lvc import make_problem
    from macaw_amd.problem import load_problem
    make_problem("t1", str(tmp_path))
    inp = (tmp_path / "t1.inp").read_text()
    inp = re.sub(r"integration = CMF2\n[^\n]*\n[^\n]*\n", f"integration = {name}\n", inp)
    (tmp_path / "v.inp").write_text(inp)
    prob = load_problem(str(tmp_path / "v.inp"))
    assert prob.int_type == method
    outs, pops = RUN.run(str(tmp_path / "v.inp"), str(tmp_path / "o"))
    e = Engine(prob)
    ref, rp, _ = e.propagate_vmf(method, prob.tinit, prob.tfinal)
    e.close()
    assert len(ref) >= len(outs)
    for a, b in zip(outs, ref):
        assert abs(a.time - b.time) < 1e-9
        assert abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im)) < 1e-12


FORTRAN_EXPECT = """
program expfmt
  implicit none
  integer :: i, n, k
  double precision :: t
  double complex :: e(3)
  read(*,*) n
  do i = 1, n
     read(*,*) t, (e(k), k = 1, 3)
     write(*,'(1x,f8.2,2x,800(3x, f14.9,1x,f14.9))') t, e
  end do
end program expfmt
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_expectation_line_matches_fortran_edit_descriptor(tmp_path):
    """The `expectation` file line of output.f90:497 (three operators)."""
    src = tmp_path / "expfmt.f90"
    src.write_text(FORTRAN_EXPECT)
    exe = tmp_path / "expfmt"
    subprocess.check_call([FLANG, "-O0", "-o", str(exe), str(src)], cwd=str(tmp_path))
    rng = np.random.default_rng(11)
    rows = [(0.0, [1.0 + 0j, -0.5 + 1e-12j, 0j])]
    for _ in range(15):
        rows.append((float(rng.uniform(0, 300)), [complex(*rng.normal(scale=3.0, size=2)) for _ in range(3)]))
    text = f"{len(rows)}\n" + "".join(
        f"{t!r} " + " ".join(f"({v.real!r},{v.imag!r})" for v in vals) + "\n" for t, vals in rows)
    out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.splitlines()
    for k, (t, vals) in enumerate(rows):
        assert RUN.expectation_line(t / AU2FS, vals) == out[k], (k, RUN.expectation_line(t / AU2FS, vals), out[k])


def test_expect_keyword_reads_the_operator_files(tmp_path):
    from macaw_amd.lvc import make_problem
    from macaw_amd.problem import load_problem
    make_problem("t1", str(tmp_path))
    (tmp_path / "q1.op").write_text("# <q1^2>\n 1.0  |  1  |1  q^2\n")
    inp = (tmp_path / "t1.inp").read_text()
    inp = inp.replace("Hamiltonian = ", "expect = q1.op\nexpect = t1.op\nHamiltonian = ")
    (tmp_path / "e.inp").write_text(inp)
    prob = load_problem(str(tmp_path / "e.inp"))
    assert [f for f, _ in prob.expect] == ["q1.op", "t1.op"]
    q1 = prob.expect[0][1]
    assert len(q1) == 2 and all(t.ops == {1: 2} for t in q1)   # "1" = every diagonal state
    assert len(prob.expect[1][1]) == prob.nHam


@pytest.mark.gpu
def test_expectation_values_match_oracle(tmp_path):
    """<Psi|O|Psi> of the standalone run's expectation operators against the
    oracle's H.A with the operator (grid modes, and Gaussian modes through
    B = D^H S A); the Hamiltonian as an expectation operator gives the
    energy of the output record (energyMF / norm)."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle_lib import Oracle
    from macaw_amd.engine import Engine
    from macaw_amd.lvc import make_problem
    from macaw_amd.problem import load_problem, with_operator
    for name in ("t2", "g1"):
        d = tmp_path / name
        d.mkdir()
        make_problem(name, str(d))
        (d / "q.op").write_text("# <q1 q2> + <q1^2>\n 1.0 | 1 |1 q |2 q\n 0.5 | 1 |1 q^2\n")
        inp = (d / f"{name}.inp").read_text().replace("Hamiltonian = ", f"expect = q.op\nexpect = {name}.op\nHamiltonian = ")
        (d / "e.inp").write_text(inp)
        prob = load_problem(str(d / "e.inp"))
        e = Engine(prob)
        e.propagate(0.0, 1.0 / AU2FS)
        A, psi = e.get_state()
        out, _ = e.observables(1.0 / AU2FS)
        vals = RUN.Expectations(prob)(A, psi)
        for (f, terms), v in zip(prob.expect, vals):
            o = Oracle(with_operator(prob, terms))
            o.set_state(A, psi)
            if name.startswith("g"):
                o.prepare_gwp()
                _, D = o.gwp_overlaps()
                B = o.a_to_b(A, D)
            else:
                o.h1mode()
                B = A
            ref = complex(np.vdot(B, o.apply_h(B)))
            assert abs(v - ref) < 1e-10 * max(1.0, abs(ref)), (name, f, v, ref)
            assert abs(v.imag) < 1e-10 * max(1.0, abs(v))
        assert abs(vals[1].real / out.norm - out.energy) < 1e-9 * abs(out.energy), (vals[1], out.energy, out.norm)
        e.close()
