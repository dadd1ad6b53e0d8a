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
