"""The batched workload at its configured size: 256 independent evolving
graphs, vertex counts log-uniform in [2,000, 200,000], average degree drawn
from {3, 5, 7}, 10 batches (90/1/9) each, maxIterations 100,000.  The batched
runner (one propagation launch per batch index, captured HIP graphs) must
give every graph the labels and per-batch iteration / update counts of the
plain one-graph-at-a-time loop, bit for bit; and, at a reduced iteration cap
that the CPU finishes in seconds, the oracle's labels for eight graphs
including the largest and the smallest."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import labelpropcuda_amd as L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def c4():
    graphs = L.batched_workload(256, 10, 42)
    sizes = [g.numNodes for g in graphs]
    assert len(graphs) == 256 and min(sizes) >= 2000 and max(sizes) <= 200000
    return graphs


def plain_loop(graphs, max_iter, workers=8):
    """Each graph in its own context, batch after batch (the plain loop);
    contexts are independent, so several run side by side on host threads
    (each on its own stream): same results, less wall time."""
    def one(inp):
        with L.Context(inp, L.make_config(inp, maxIterations=max_iter)) as ctx:
            ctx.run_all()
            ctx.synchronize()
            return ctx.labels(), [(s["iterations"], s["updates"])
                                  for s in (ctx.stats(b) for b in range(inp.numBatches))]
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(one, graphs))


def multi_results(graphs, max_iter, **tuning):
    cfgs = [L.make_config(g, maxIterations=max_iter) for g in graphs]
    with L.Multi(graphs, cfgs) as m:
        if tuning:
            m.set_tuning(**tuning)
        m.run_all()
        m.synchronize()
        return [(m.labels(g), [(s["iterations"], s["updates"])
                               for s in (m.stats(g, b) for b in range(inp.numBatches))])
                for g, inp in enumerate(graphs)]


@pytest.mark.timeout(900)
def test_c4_batched_matches_plain_for_all_256_graphs(gpu_available, c4):
    plain = plain_loop(c4, 100000)
    batched = multi_results(c4, 100000)
    total = 0
    for g in range(len(c4)):
        assert np.array_equal(bits(batched[g][0]), bits(plain[g][0])), f"graph {g}: labels"
        assert batched[g][1] == plain[g][1], f"graph {g}: per-batch iterations / updates"
        total += sum(u for _, u in plain[g][1])
    print(f"C4: 256 graphs, {total} affected-vertex updates, identical in both forms")


@pytest.mark.timeout(900)
def test_c4_batched_matches_oracle_on_eight_graphs(gpu_available, c4):
    cap = 3000
    batched = multi_results(c4, cap)
    sizes = np.array([g.numNodes for g in c4])
    rng = np.random.default_rng(4)
    pick = [int(np.argmax(sizes)), int(np.argmin(sizes))]
    pick += [int(x) for x in rng.choice(np.setdiff1d(np.arange(len(c4)), pick), 6,
                                        replace=False)]
    for g in pick:
        orc = Oracle(c4[g], maxIterations=cap)
        for b in range(c4[g].numBatches):
            orc.run_batch(b)
            st = orc.stats()
            assert (st["iterations"], st["updates"]) == batched[g][1][b], f"graph {g} batch {b}"
        assert np.array_equal(bits(batched[g][0]), bits(orc.labels())), f"graph {g}: labels"
