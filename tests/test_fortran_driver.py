"""The Fortran side of the boundary: macaw_gpu.f90 (ISO_C_BINDING) and a
Fortran driver that runs the CMF2 loop on the engine and writes Macaw's
`auto` / `spop` files."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from macaw_amd.fortran_io import read_auto, write_descriptor
from macaw_amd.lvc import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "macaw_amd", "fortran")
EXE = os.path.join(ROOT, "macaw_amd", "_build", "fortran", "drive_cmf2")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def build_driver():
    if not os.path.exists(os.path.join(ROOT, "macaw_amd", "libmacaw_gpu.so")):
        from macaw_amd.build import build
        build()
    subprocess.check_call(["make", "-s", "-C", FDIR])
    return EXE


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_binding_builds_and_links():
    exe = build_driver()
    assert os.path.exists(exe)
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libmacaw_gpu.so" in out and "not found" not in out


def test_descriptor_file_layout(tmp_path):
    prob = make_problem("t1", str(tmp_path))
    path = str(tmp_path / "t1.bin")
    write_descriptor(prob, path, 10.0)
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:72], dtype=np.int64)
    assert list(hdr[:8]) == [2, 3, 3, 1, 8, prob.nHam, prob.nAConftot, prob.dimpsi]
    n_i64 = 3 + 3 + 3 + 6 + 3 + 2 * prob.nHam + 3 * prob.nHam
    n_f64 = 8 * 3 + 3 * 8 * 8 * 3 + 2 * prob.nHam + 2 * prob.nAConftot + 2 * prob.dimpsi
    assert len(raw) == 72 + 48 + 8 * n_i64 + 8 * n_f64


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_driver_matches_oracle(tmp_path):
    from oracle_lib import Oracle
    exe = build_driver()
    prob = make_problem("t1", str(tmp_path))
    tf = 3.0 / 0.02418884326505
    write_descriptor(prob, str(tmp_path / "t1.bin"), tf)
    subprocess.check_call([exe, str(tmp_path / "t1.bin")], cwd=str(tmp_path), timeout=300)
    auto = read_auto(str(tmp_path / "auto"))
    outs, pops, _ = Oracle(prob).propagate(0.0, tf)
    assert auto.shape[0] == len(outs)
    for row, o in zip(auto, outs):
        assert abs(row[0] - o.time * 0.02418884326505) < 1e-4
        assert abs(row[1] - o.auto_re) < 1e-8 and abs(row[2] - o.auto_im) < 1e-8
    spop = np.loadtxt(str(tmp_path / "spop"), comments="#")
    assert np.abs(spop[:, 1:] - pops).max() < 2e-7   # f10.7 output


REF_SRC = "/root/reference/source"
FFLAGS = ["-fdefault-integer-8", "-fdefault-real-8", "-fdefault-double-8"]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="Macaw sources not present")
def test_cmf2_base_dropin_compiles_against_macaw_modules(tmp_path):
    """The drop-in module of INTEGRATION.md, committed as
    macaw_amd/fortran/cmf2_base_dropin.f90, compiled against Macaw's own
    sysparam / globals / psidef (built here from the reference sources, as
    they are) and declaration mirrors of dvr / hamilton with the reference's
    attributes, then linked into a propagate_wf-shaped program with the
    engine library.  Module arrays without TARGET would fail c_loc here."""
    if not os.path.exists(os.path.join(ROOT, "macaw_amd", "libmacaw_gpu.so")):
        from macaw_amd.build import build
        build()
    d = str(tmp_path)
    srcs = [os.path.join(REF_SRC, f) for f in ("sysparam.f90", "globals.f90", "psidef.f90")]
    srcs += [os.path.join(ROOT, "tests", "fortran", "dvr_hamilton_decl.f90"),
             os.path.join(FDIR, "macaw_gpu.f90"), os.path.join(FDIR, "cmf2_base_dropin.f90"),
             os.path.join(ROOT, "tests", "fortran", "use_dropin.f90")]
    objs = []
    for src in srcs:
        obj = os.path.join(d, os.path.basename(src) + ".o")
        r = subprocess.run([FLANG, *FFLAGS, "-J", d, "-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, (src, r.stderr)
        objs.append(obj)
    exe = os.path.join(d, "use_dropin")
    r = subprocess.run([FLANG, "-o", exe, *objs, "-L" + os.path.join(ROOT, "macaw_amd"), "-lmacaw_gpu",
                        "-Wl,-rpath," + os.path.join(ROOT, "macaw_amd")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    syms = subprocess.run(["nm", objs[-2]], capture_output=True, text=True).stdout
    for name in ("mg_create", "mg_set_start", "mg_set_state", "mg_cmf2_step", "mg_get_state"):
        assert " U " + name in syms


def test_every_exported_entry_point_has_a_fortran_binding():
    """include/macaw_gpu.h and macaw_amd/fortran/macaw_gpu.f90 declare the
    same entry points, and every bound name is public."""
    import re
    hdr = open(os.path.join(ROOT, "include", "macaw_gpu.h")).read()
    exported = set(re.findall(r"^(?:int|int64_t|const char \*)\s*(mg_\w+)\(", hdr, re.M))
    f90 = open(os.path.join(FDIR, "macaw_gpu.f90")).read()
    bound = set(re.findall(r"bind\(C, name='(mg_\w+)'\)", f90))
    public = set()
    for line in re.findall(r"^\s*public ::(.*)$", f90, re.M):
        public |= {x.strip() for x in line.split(",")}
    assert exported and exported == bound, (exported - bound, bound - exported)
    assert bound <= public, bound - public
