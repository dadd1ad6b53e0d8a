"""A long run split over processes: a run stopped and resumed from its saved
state (macaw_amd.trajectory --resume) writes the same output records as the
uninterrupted run."""
import numpy as np
import pytest

from macaw_amd import trajectory as T

pytestmark = pytest.mark.gpu


def test_resumed_run_matches_uninterrupted(tmp_path):
    one, two = str(tmp_path / "one"), str(tmp_path / "two")
    common = ["--workload", "t1", "--precision", "0", "--controller", "classical", "--t-end", "3.0"]
    T.main(["run", *common, "--out", one])
    T.main(["run", *common, "--out", two, "--max-intervals", "2"])
    part = T.read_outputs(two + "/outputs.txt")
    full = T.read_outputs(one + "/outputs.txt")
    assert 0 < len(part) < len(full)
    T.main(["run", *common, "--out", two, "--resume"])
    res = T.read_outputs(two + "/outputs.txt")
    assert res.shape == full.shape
    # same step size carried over, same classical controller: the same arithmetic
    assert np.abs(res - full).max() <= 1e-12
    a1, a2 = T.read_attempts(one + "/attempts.txt"), T.read_attempts(two + "/attempts.txt")
    assert [(a.time, a.h, a.accepted) for a in a1] == [(a.time, a.h, a.accepted) for a in a2]
