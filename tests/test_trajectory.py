"""macaw_amd.trajectory: the on-disk record of long runs and their
comparison (host logic; the runs themselves need the GPU)."""
import json
import os

import numpy as np

from macaw_amd import trajectory as T


def _write_run(d, times, c, precision, replay=None, fs_per_s=0.1):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "outputs.txt"), "w") as fh:
        fh.write("# header\n")
        for t, z in zip(times, c):
            fh.write(f"{t:.10f} {z.real:.17e} {z.imag:.17e} 1.0 0.01 0.4 0.6\n")
    with open(os.path.join(d, "attempts.txt"), "w") as fh:
        fh.write("# header\n")
        fh.write(f"{0.0!r} {1.5!r} {1e-9!r} {2e-9!r} 1 4 5\n")
        fh.write(f"{1.5!r} {2.0!r} {1e-8!r} {3e-8!r} 0 6 7\n")
    json.dump({"precision": precision, "replay": replay, "t_reached_fs": times[-1], "fs_per_s": fs_per_s,
               "attempts": 2}, open(os.path.join(d, "meta.json"), "w"))


def test_compare_reports_the_largest_differences(tmp_path, capsys):
    t = np.arange(0, 20.5, 0.5)
    c = np.exp(-1j * t / 3)
    _write_run(str(tmp_path / "ref"), t, c, 0)
    bump = np.zeros_like(c)
    bump[30] = 3e-7j
    _write_run(str(tmp_path / "mix"), t, c + bump, 1, replay="ref/attempts.txt")
    md = str(tmp_path / "cmp.md")
    T.main(["compare", "--ref", str(tmp_path / "ref"), "--runs", str(tmp_path / "mix"), "--md", md])
    text = open(md).read()
    assert "3.00e-07" in text and "| 1 | replayed |" in text
    assert "20: 3.00e-07" in text and "10: 0.00e+00" in text


def test_attempt_log_round_trip(tmp_path):
    _write_run(str(tmp_path / "r"), [0.0, 0.5], [1.0, 0.9], 0)
    a = T.read_attempts(str(tmp_path / "r" / "attempts.txt"))
    assert len(a) == 2 and a[1].accepted == 0 and a[1].nkrylov2 == 7 and a[0].h == 1.5
