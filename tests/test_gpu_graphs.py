"""HIP-graph replay of the cached launch plans and of whole CMF2 attempts.

Graphs hold the same kernels with the same arguments as the eager launches,
so every replay must give bit-identical results: H.A replayed several times
with eager work and host synchronisation between the replays, and whole
CMF2 propagations (grid modes FP64 and mixed, several mode groups; Gaussian
modes) with graphs on and off.
"""
import numpy as np
import pytest

from macaw_amd.lvc import make_problem

gpu = pytest.mark.gpu
AU2FS = 0.02418884326505


def engine(prob, precision=0, graphs=True):
    from macaw_amd.engine import Engine
    e = Engine(prob, precision=precision)
    e.set_graphs(graphs)
    return e


@gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_apply_h_replays_match_eager(precision, tmp_path):
    prob = make_problem("t5", str(tmp_path))
    rng = np.random.default_rng(4)
    eg, ee = engine(prob, precision, True), engine(prob, precision, False)
    for e in (eg, ee):
        e.h1mode()
    for rep in range(5):
        B = rng.normal(size=prob.nAConftot) + 1j * rng.normal(size=prob.nAConftot)
        Hg = eg.apply_h(B)
        eg.rho()                       # eager work between replays
        He = ee.apply_h(B)
        assert np.array_equal(Hg, He), f"replay {rep}"


@gpu
@pytest.mark.parametrize("name,precision,tf_fs,attempt_graph,fork",
                         [("t5", 0, 1.0, 1, 1), ("t5", 1, 1.0, 1, 1), ("t2", 0, 2.0, 1, 1), ("g3", 0, 0.3, 1, 1),
                          ("t5", 1, 1.0, 1, 0), ("t5", 1, 1.0, 0, 1), ("g3", 0, 0.3, 0, 1)])
def test_propagation_with_graphs_is_bit_identical(name, precision, tf_fs, attempt_graph, fork, tmp_path,
                                                  monkeypatch):
    """Whole attempts as one graph (MG_ATTEMPT_GRAPH=1), with and without the
    second stream for the SPF half steps, and the per-plan graphs alone."""
    prob = make_problem(name, str(tmp_path))
    tf = tf_fs / AU2FS
    monkeypatch.setenv("MG_ATTEMPT_GRAPH", str(attempt_graph))
    monkeypatch.setenv("MG_ATTEMPT_FORK", str(fork))
    og, pg, ag = engine(prob, precision, True).propagate(0.0, tf)
    oe, pe, ae = engine(prob, precision, False).propagate(0.0, tf)
    assert len(ag) == len(ae)
    for a, b in zip(og, oe):
        assert (a.auto_re, a.auto_im, a.norm, a.energy) == (b.auto_re, b.auto_im, b.norm, b.energy)
    assert np.array_equal(pg, pe)


@gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_krylov_build_outgrowing_the_graph_is_repeated(precision, tmp_path, monkeypatch):
    """A CMF2 attempt runs as one graph with a fixed number of Lanczos
    iterations per Krylov build; a build that needs more is repeated with the
    reference's limit.  Capping the graph at 2 iterations forces that on
    every attempt: the trajectory equals the uncapped one bit for bit."""
    prob = make_problem("t5", str(tmp_path))
    tf = 1.0 / AU2FS
    monkeypatch.setenv("MG_ATTEMPT_GRAPH", "1")
    oa, pa, aa = engine(prob, precision, True).propagate(0.0, tf)
    monkeypatch.setenv("MG_ATTEMPT_KMAX", "2")
    ob, pb, ab = engine(prob, precision, True).propagate(0.0, tf)
    assert [(a.h, a.accepted, a.nkrylov1, a.nkrylov2) for a in aa] == \
        [(a.h, a.accepted, a.nkrylov1, a.nkrylov2) for a in ab]
    for a, b in zip(oa, ob):
        assert (a.auto_re, a.auto_im, a.norm, a.energy) == (b.auto_re, b.auto_im, b.norm, b.energy)
    assert np.array_equal(pa, pb)
