"""mg_mlp_forward_noalloc (no counterpart in the reference): the learned
controller's forward pass with its device memory in a fixed file-scope array
instead of allocations, for callers that cannot allocate.  Checked through
the C-ABI against mg_mlp_forward on the same inputs, on two calls in a row
(the second smaller, so stale contents of the shared array would show), and
its size limit."""
import os

import numpy as np
import pytest

from macaw_amd.lvc import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_mlp_forward_noalloc_matches_mg_mlp_forward(tmp_path):
    from macaw_amd.engine import Engine, EngineError
    print("no-alloc: t1 engine, shipped step-controller network, 3000 then 77 rows", flush=True)
    e = Engine(make_problem("t1", str(tmp_path)))
    params = np.load(os.path.join(ROOT, "macaw_amd", "step_controller.npy")).astype(np.float32)
    rng = np.random.default_rng(7)
    for n in (3000, 77):
        X = rng.normal(size=(n, 12)).astype(np.float32)
        y_ref = e.mlp_forward(params, X)
        y = e.mlp_forward_noalloc(params, X)
        assert y.shape == (n,) and np.all(np.isfinite(y))
        assert np.array_equal(y, y_ref), np.abs(y - y_ref).max()
    with pytest.raises(EngineError, match="between 1 and 4096 rows"):
        e.mlp_forward_noalloc(params, np.zeros((4097, 12), np.float32))
    # the context is still usable and the array still serves a full batch
    X = rng.normal(size=(4096, 12)).astype(np.float32)
    assert np.array_equal(e.mlp_forward_noalloc(params, X), e.mlp_forward(params, X))
    e.close()
