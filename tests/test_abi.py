"""CPU-side checks of the C-ABI boundary: the HIP library loads and exports
every entry point declared in include/*.h (no device calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "macaw_amd", "libmacaw_gpu.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|int64_t|const char \*|void)\s*\*?\s*(mg_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from macaw_amd.build import build
        build()
    return ctypes.CDLL(LIB)


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("mg_create", "mg_apply_h", "mg_cmf2_step", "mg_propagate", "mg_calc_meanfield"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_header():
    from macaw_amd.abi import MgAttempt, MgOutput, MgProblem
    assert ctypes.sizeof(MgOutput) == 5 * 8
    assert ctypes.sizeof(MgAttempt) == 7 * 8
    # 4 int64, 5 pointers, ngpmax, 4 pointers, nham, 4 pointers, 5 doubles, int64,
    # then the Gaussian-mode fields: 2 pointers, 2 doubles; then the
    # parametrised operators: nopparam and the gridoper pointer
    assert ctypes.sizeof(MgProblem) == 8 * (4 + 5 + 1 + 4 + 1 + 4 + 5 + 1 + 2 + 2 + 2)
    assert MgProblem.nopparam.offset == 8 * 29 and MgProblem.gridoper.offset == 8 * 30


def test_input_deck_roundtrip(tmp_path):
    from macaw_amd.lvc import SPECS, draw, make_problem
    prob = make_problem("t2", str(tmp_path))
    omega, kappa, gamma, couplings = draw(SPECS["t2"])
    # 2 terms per mode for every state, 1 per state gap, kappa, gamma, 2 per coupling
    S, f = 3, 3
    assert prob.nHam == 2 * f * S + S + S * f + S * f + 2 * len(couplings)
    assert prob.nAConftot == 27 + 64 + 8
    assert abs(prob.ae[0].real + omega[0] / 2 / 27.2113961) < 1e-14
