"""Learned CMF2 step controller.

CPU: the numpy reference's gradient against finite differences (this pins
the reference the kernels are checked against), BF16 rounding, the
training-pair construction and the error-exponent estimate.  GPU (through
the C-ABI): forward 1e-5 and gradient 1e-4 against the reference, training
that lowers the loss on a known function, and a propagation under a model
whose accepted steps still pass the unchanged CMF2 criterion.
"""
import math
import types

import numpy as np
import pytest

import mlp_ref as R
from macaw_amd.controller import Y_MAX, Y_MIN, error_exponent, training_pairs

gpu = pytest.mark.gpu


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = R.random_params(rng).astype(np.float64)
    X = rng.normal(size=(7, 12))
    t = rng.normal(size=7)
    _, g = R.loss_grad(p, X, t, round_weights=False)
    for idx in [R.oW1 + 5, R.oB1 + 3, R.oW2 + 700, R.oB2 + 10, R.oW3 + 40, R.oB3]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        lp, _ = R.loss_grad(pp, X, t, round_weights=False)
        lm, _ = R.loss_grad(pm, X, t, round_weights=False)
        assert abs((lp - lm) / (2 * h) - g[idx]) < 1e-6 * max(1.0, abs(g[idx]))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -3.14159, 1e-3], np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                      # tie to even
    assert r[2] == np.float32(1.0 + 2 ** -6)                # tie rounds up to even
    assert abs(r[3] - x[3]) <= 2 ** -8 * abs(x[3])
    assert (r.view(np.uint32) & 0xFFFF).max() == 0


def att(h, ea, acc, eps=1e-7):
    return types.SimpleNamespace(h=h, err_A=ea * eps, err_psi=0.0, accepted=acc)


def test_error_exponent_from_rejected_retries():
    # err ~ h^3: rejected h=2 (err 8), retried h=1 (err 1)
    atts = [att(1.0, 0.5, 1), att(2.0, 8.0, 0), att(1.0, 1.0, 1)]
    assert abs(error_exponent(atts, 1e-7) - 3.0) < 1e-12
    assert error_exponent([att(1.0, 0.5, 1)], 1e-7) == 2.0


def test_training_pairs_targets():
    atts = [att(1.0, 0.5, 1), att(1.5, 0.81 * 4, 0), att(0.75, 0.81, 1), att(1.0, 0.2, 1)]
    feats = np.arange(36, dtype=np.float32).reshape(3, 12)
    X, y, p = training_pairs(feats, [1.0, 0.75, 1.0], atts, 1e-7, exponent=2.0)
    # after step 0 (h=1): next attempt h=1.5, err ratio 3.24 -> h* = 1.5 * 0.5 = 0.75
    assert abs(y[0] - math.log(0.75)) < 1e-6
    # after step 1 (h=0.75): next h=1, err 0.2 -> h* = 1 * sqrt(0.81/0.2) > 2 h: clipped
    assert abs(y[1] - min(Y_MAX, math.log(math.sqrt(0.81 / 0.2) / 0.75))) < 1e-6
    assert len(y) == 2 and (X[1] == feats[1]).all()
    assert Y_MIN <= y.min() and y.max() <= Y_MAX


@gpu
def test_forward_and_gradient_match_reference(tmp_path):
    from macaw_amd.engine import Engine
    from macaw_amd.lvc import make_problem
    e = Engine(make_problem("t1", str(tmp_path)))
    rng = np.random.default_rng(2)
    p = R.random_params(rng)
    X = rng.normal(size=(300, 12)).astype(np.float32)
    t = rng.normal(size=300).astype(np.float32)
    y_ref, _ = R.forward(p, X)
    y = e.mlp_forward(p, X)
    assert np.abs(y - y_ref).max() < 1e-5 * max(1.0, np.abs(y_ref).max())
    g, loss = e.mlp_grad(p, X, t)
    l_ref, g_ref = R.loss_grad(p, X, t)
    assert abs(loss - l_ref) < 1e-5 * l_ref
    assert np.abs(g - g_ref).max() < 1e-4 * max(1.0, np.abs(g_ref).max())


@gpu
def test_training_fits_a_known_function(tmp_path):
    from macaw_amd.engine import Engine
    from macaw_amd.lvc import make_problem
    e = Engine(make_problem("t1", str(tmp_path)))
    rng = np.random.default_rng(3)
    X = rng.normal(size=(4000, 12)).astype(np.float32) * 2 + 1
    t = (0.3 * np.tanh(X[:, 0] - 1) - 0.2 * (X[:, 4] - 1) / 2 + 0.1 * np.sin(X[:, 8])).astype(np.float32)
    params, loss = e.mlp_train(X, t, epochs=30, batch=128, lr=2e-3, seed=5)
    assert loss < 0.05 * np.var(t)
    y = e.mlp_forward(params, X)
    assert np.mean((y - t) ** 2) == pytest.approx(loss, rel=1e-3)


@gpu
def test_learned_controller_keeps_the_acceptance_criterion(tmp_path):
    """Train on a logged classical run of t1 and propagate with the model:
    every accepted attempt still has err <= eps_int, the model proposed the
    steps, and C(t) stays within the CMF2 tolerance of the classical run."""
    from macaw_amd.engine import Engine
    from macaw_amd.lvc import make_problem
    prob = make_problem("t1", str(tmp_path))
    tf = 3.0 / 0.02418884326505
    e = Engine(prob)
    e.log_step_features(True)
    outs_c, _, atts_c = e.propagate(0.0, tf)
    f, h = e.step_features()
    X, y, p = training_pairs(f, h, atts_c, prob.eps_int)
    assert len(y) > 20 and 1.0 <= p <= 4.0
    params, loss = e.mlp_train(X, y, epochs=60, batch=32, lr=2e-3, seed=1)
    e2 = Engine(prob)
    e2.set_step_model(params)
    outs_m, _, atts_m = e2.propagate(0.0, tf)
    prop, fb = e2.controller_stats()
    assert prop > 10 and fb == 0
    for a in atts_m:
        if a.accepted:
            assert max(a.err_A, a.err_psi) <= prob.eps_int
    dc = max(abs(complex(a.auto_re, a.auto_im) - complex(b.auto_re, b.auto_im)) for a, b in zip(outs_c, outs_m))
    print(f"classical {len(atts_c)} attempts ({sum(not a.accepted for a in atts_c)} rejected); learned "
          f"{len(atts_m)} ({sum(not a.accepted for a in atts_m)} rejected); |dC| {dc:.2e}")
    assert dc < 1e-6


@gpu
def test_shipped_controller_cuts_rejections_on_a_fresh_seed(tmp_path):
    """The committed network (trained on logged classical runs of the 2- and
    4-state grid workloads and the G-MCTDH workload, other seeds) against the
    classical controller on a seed it never saw: fewer rejected attempts,
    and C(t) as close to a run at half the tolerance as the classical run."""
    import os
    from macaw_amd.controller import evaluate, seeded_spec
    from macaw_amd.lvc import make_problem
    params = np.load(os.path.join(os.path.dirname(__file__), "..", "macaw_amd", "step_controller.npy"))
    prob = make_problem(seeded_spec("c1", 4242), str(tmp_path))
    res = evaluate(prob, params, 4.0)
    c, m = res["classical"], res["learned"]
    assert m["model_fallbacks"] == 0 and m["model_proposals"] > 0
    assert m["rejected"] <= c["rejected"] // 2 + 1
    assert m["max_dC_vs_reference"] <= max(2.0 * c["max_dC_vs_reference"], 1e-7)
