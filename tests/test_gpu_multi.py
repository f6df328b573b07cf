"""The multi-graph runner (one propagation launch per batch index for many
graphs) against one context per graph and against the oracle, through the
C-ABI.  Labels are compared bit for bit; iteration and update counts
exactly."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
F32 = np.float32


def bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


@pytest.fixture(scope="module")
def small_graphs():
    """Ten graphs of 1,000-20,000 vertices, degrees 3/5/7; two of them have
    fewer batches than the rest."""
    rng = np.random.default_rng(7)
    out = []
    for g in range(10):
        n = int(rng.integers(1000, 20000))
        d = int(rng.choice([3, 5, 7]))
        b = 4 if g in (3, 7) else 6
        out.append(L.generate(n, d, b, 100 + g))
    return out


def plain_results(graphs, maxIter):
    res = []
    for inp in graphs:
        with L.Context(inp, L.make_config(inp, maxIterations=maxIter)) as ctx:
            ctx.run_all()
            res.append((ctx.labels(), [ctx.stats(b) for b in range(inp.numBatches)]))
    return res


def check_multi(m, graphs, plain):
    for g, inp in enumerate(graphs):
        lab, st = plain[g]
        got = m.labels(g)
        assert np.array_equal(bits(got), bits(lab)), f"graph {g}: labels differ"
        for b in range(inp.numBatches):
            s = m.stats(g, b)
            assert (s["iterations"], s["updates"], s["frontier0"], s["numMarked"]) == \
                (st[b]["iterations"], st[b]["updates"], st[b]["frontier0"],
                 st[b]["numMarked"]), f"graph {g} batch {b}"


@pytest.mark.parametrize("shape", [3, 1, 5])
def test_multi_matches_one_context_per_graph(gpu_available, small_graphs, shape):
    plain = plain_results(small_graphs, 3000)
    cfgs = [L.make_config(g, maxIterations=3000) for g in small_graphs]
    with L.Multi(small_graphs, cfgs) as m:
        m.set_tuning(localShape=shape)
        m.run_all()
        m.synchronize()
        check_multi(m, small_graphs, plain)
        # second step replays the captured batches
        m.run_all()
        m.synchronize()
        check_multi(m, small_graphs, plain)


@pytest.mark.parametrize("threshold,graphs_mode", [(-1, 1), (5, 2), (100, 2)])
def test_multi_engine_modes(gpu_available, small_graphs, threshold, graphs_mode):
    """HBM iterations only, a tiny LDS threshold (many hand-overs), and the
    enqueue-every-time mode."""
    plain = plain_results(small_graphs, 3000)
    cfgs = [L.make_config(g, maxIterations=3000) for g in small_graphs]
    with L.Multi(small_graphs, cfgs) as m:
        m.set_tuning(thresholdPercent=threshold, useGraphs=graphs_mode)
        m.run_all()
        m.synchronize()
        check_multi(m, small_graphs, plain)


@pytest.mark.parametrize("heavy,graphs_mode", [(1, 1), (100, 1), (100, 2), (-1, 1)])
def test_multi_hand_back(gpu_available, small_graphs, heavy, graphs_mode):
    """Graphs whose frontier outgrows `heavy` leave the shared launch and
    finish on the single-graph loop (1: almost every graph at its first
    iteration; -1: never); two steps, so replays are covered too."""
    plain = plain_results(small_graphs, 3000)
    cfgs = [L.make_config(g, maxIterations=3000) for g in small_graphs]
    with L.Multi(small_graphs, cfgs) as m:
        m.set_tuning(heavyFrontier=heavy, useGraphs=graphs_mode)
        assert m.tuning()["heavyFrontier"] == heavy
        for _ in range(2):
            m.run_all()
            m.synchronize()
            check_multi(m, small_graphs, plain)


def test_multi_matches_oracle(gpu_available):
    graphs = [L.generate(3000, 5, 5, 11), L.generate(6000, 3, 5, 12),
              L.generate(2500, 7, 5, 13)]
    with L.Multi(graphs) as m:
        m.reset()
        orcs = [Oracle(g) for g in graphs]
        for b in range(5):
            m.run_batch(b)
            m.synchronize()
            for g, o in enumerate(orcs):
                o.run_batch(b)
                assert np.array_equal(bits(m.labels(g)), bits(o.labels())), f"graph {g} b {b}"
                s, os_ = m.stats(g, b), o.stats()
                assert (s["iterations"], s["updates"]) == (os_["iterations"], os_["updates"])


def test_multi_long_rows(gpu_available):
    """Degree 40: rows longer than the lane paths in both engines."""
    graphs = [L.generate(2500, 40, 4, 21), L.generate(4000, 25, 4, 22)]
    plain = plain_results(graphs, 2000)
    cfgs = [L.make_config(g, maxIterations=2000) for g in graphs]
    for th in (100, -1):
        with L.Multi(graphs, cfgs) as m:
            m.set_tuning(thresholdPercent=th)
            m.run_all()
            m.synchronize()
            check_multi(m, graphs, plain)
