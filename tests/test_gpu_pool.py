"""The runner's scratch pool on the GPU: batches draw their per-batch
scratch from one shared allocation.  The HBM budget is forced from the host
(a pool sized against the batches' own scratch needs), and every admission
decision it forces -- defer, reject, run -- must leave the labels and the
per-batch counts identical to each tenant run alone."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import _lib

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def pair():
    """two tenants whose second batches are of similar size"""
    return [L.generate_stream(20000, 5, [0, 3000, 500, 2000], 501),
            L.generate_stream(16000, 6, [0, 2500, 800, 1500], 502)]


def alone(tenants, maxIter=3000):
    res = []
    for inp in tenants:
        with L.Context(inp, L.make_config(inp, maxIterations=maxIter)) as ctx:
            ctx.run_all()
            res.append((ctx.labels(), [ctx.stats(b) for b in range(inp.numBatches)]))
    return res


def same_as_alone(m, tenants, ref):
    for t, inp in enumerate(tenants):
        assert np.array_equal(bits(m.labels(t)), bits(ref[t][0])), f"tenant {t}"
        for b in range(inp.numBatches):
            s = m.stats(t, b)
            assert (s["iterations"], s["updates"]) == \
                (ref[t][1][b]["iterations"], ref[t][1][b]["updates"]), f"tenant {t} b {b}"


def test_budget_below_two_batches_defers_the_second(gpu_available, pair):
    """A pool that holds either tenant's next batch but not both: the two
    requests arrive together, so one is admitted and the other deferred
    until the first returns its scratch."""
    ref = alone(pair)
    cfgs = [L.make_config(t, maxIterations=3000) for t in pair]
    with L.Multi(pair, cfgs) as m:
        m.reset()
        m.run_next([0, 1])
        for b in (1, 2, 3):
            need = [m.batch_scratch_bytes(0, b), m.batch_scratch_bytes(1, b)]
            budget = max(need) + 256
            assert budget < sum(need)
            print(f"batch {b}: needs {need}, pool {budget}", flush=True)
            m.set_scratch_pool(budget)  # fresh counters for this batch
            done, steps = m.serve([(0, 0.0), (1, 0.0)], maxSlots=8, hbmBudgetBytes=budget)
            st = m.pool_stats()
            assert st["deferrals"] >= 1, st
            assert st["peak"] <= st["capacity"] and st["inUse"] == 0
            assert st["peak"] == max(need)  # never both at once
        m.synchronize()
        same_as_alone(m, pair, ref)


def test_budget_below_one_batch_is_rejected(gpu_available, pair):
    """A pool smaller than a batch's scratch cannot ever run it: the call
    fails with DLP_ERR_NOMEM, and the runner works once the pool is large
    enough (no lease is left behind, no state advanced)."""
    ref = alone(pair)
    cfgs = [L.make_config(t, maxIterations=3000) for t in pair]
    with L.Multi(pair, cfgs) as m:
        m.reset()
        m.run_next([0, 1])
        need = m.batch_scratch_bytes(0, 1)
        with pytest.raises(_lib.DynLPError) as e:
            m.serve([(0, 0.0)], maxSlots=4, hbmBudgetBytes=need - 256)
        assert e.value.code == _lib.DLP_ERR_NOMEM and "scratch" in str(e.value)
        assert m.pool_stats()["inUse"] == 0
        done, _ = m.serve([(0, 0.0), (1, 0.0)] * 3, maxSlots=4, hbmBudgetBytes=need * 4)
        m.synchronize()
        same_as_alone(m, pair, ref)


def test_lockstep_batches_from_the_pool(gpu_available, pair):
    """dlp_multi_run_batch with a pool: every graph of the batch leases at
    once; too small a pool is DLP_ERR_NOMEM before anything runs; a large
    enough one gives the per-graph-scratch results; the pool replaces the
    graphs' own scratch in the device footprint."""
    ref = alone(pair)
    cfgs = [L.make_config(t, maxIterations=3000) for t in pair]
    with L.Multi(pair, cfgs) as m:
        own = m.device_bytes()
        need0 = sum(m.batch_scratch_bytes(g, 0) for g in range(2))
        assert sum(m.batch_scratch_bytes(g, 1) for g in range(2)) > need0
        m.set_scratch_pool(need0)  # the seed batches of both graphs, not batch 1
        pooled = m.device_bytes()
        m.reset()
        m.run_batch(0)
        with pytest.raises(_lib.DynLPError) as e:
            m.run_batch(1)
        assert e.value.code == _lib.DLP_ERR_NOMEM
        assert m.pool_stats()["inUse"] == 0
        big = max(sum(m.batch_scratch_bytes(g, b) for g in range(2)) for b in range(4))
        m.set_scratch_pool(big)
        for b in range(1, 4):
            m.run_batch(b)
        m.synchronize()
        same_as_alone(m, pair, ref)
        st = m.pool_stats()
        assert st["leases"] == 6 and st["inUse"] == 0 and st["peak"] <= big
        m.set_scratch_pool(0)  # back to per-graph scratch
        assert m.device_bytes() == own
        m.run_all()
        m.synchronize()
        same_as_alone(m, pair, ref)
    assert pooled < own


def test_many_tenants_share_a_small_pool(gpu_available):
    """Twelve tenants, a pool for about three batches, staggered arrivals,
    short slices and hand-backs: labels as alone, peak within the pool."""
    rng = np.random.default_rng(9)
    tenants = []
    for t in range(12):
        n = int(rng.integers(2000, 40000))
        sizes = [0] + [int(x) for x in rng.integers(50, 3000, 4)]
        tenants.append(L.generate_stream(n, 5, sizes, 900 + t))
    ref = alone(tenants)
    cfgs = [L.make_config(t, maxIterations=3000) for t in tenants]
    with L.Multi(tenants, cfgs) as m:
        m.set_tuning(heavyFrontier=400)
        m.reset()
        m.run_next(list(range(len(tenants))))
        needs = sorted(m.batch_scratch_bytes(t, b) for t in range(12) for b in range(1, 5))
        budget = sum(needs[-3:])
        reqs = sorted([(t, float(x)) for t in range(12)
                       for x in np.sort(rng.uniform(0, 0.03, 4))], key=lambda r: r[1])
        done, steps = m.serve(reqs, maxSlots=16, hbmBudgetBytes=budget, handBackLanes=3,
                              sliceIterations=64)
        m.synchronize()
        st = m.pool_stats()
        assert st["peak"] <= budget and st["inUse"] == 0 and st["leases"] == len(reqs)
        same_as_alone(m, tenants, ref)
