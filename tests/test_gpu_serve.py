"""Serving: arrival-driven continuous batching over tenant graphs.  Every
tenant's result must be the same served alone (its own context, one batch
at a time) and in the mix, bit for bit."""
import numpy as np
import pytest

import labelpropcuda_amd as L

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def tenants():
    rng = np.random.default_rng(5)
    out = []
    for t in range(6):
        n = int(rng.integers(3000, 30000))
        nb = int(rng.integers(3, 7))
        sizes = [0] + [int(x) for x in np.rint(np.exp(rng.uniform(np.log(50), np.log(4000), nb)))]
        out.append(L.generate_stream(n, 5, sizes, 1042 + t))
    return out


def alone(tenants, maxIter):
    res = []
    for inp in tenants:
        with L.Context(inp, L.make_config(inp, maxIterations=maxIter)) as ctx:
            ctx.run_all()
            res.append((ctx.labels(), [ctx.stats(b) for b in range(inp.numBatches)]))
    return res


def requests(tenants, rng, spread):
    """Every batch after the seed batch of every tenant, arrival times
    uniform in [0, spread) seconds, sorted; a tenant's batches in order."""
    reqs = []
    for t, inp in enumerate(tenants):
        times = np.sort(rng.uniform(0, spread, inp.numBatches - 1))
        reqs += [(t, float(x)) for x in times]
    return sorted(reqs, key=lambda r: r[1])


# budget "one": a scratch pool that holds the largest batch (smaller ones
# may share it); slice (iterations, ms): 0/0.0 = the defaults, -1/-1.0 = each
# step runs its batches to the end
@pytest.mark.parametrize("slots,budget,plain,spread,heavy,lanes,slice", [
    (64, 0, False, 0.0, 0, 0, (0, 0.0)),        # everything at once: maximal mixing
    (3, 0, False, 0.05, 0, 0, (0, 0.0)),        # few slots, staggered arrivals
    (64, "one", False, 0.0, 0, 0, (0, 0.0)),    # pool for about one batch: deferrals
    (1, 0, True, 0.0, 0, 0, (0, 0.0)),          # plain: one batch at a time, single-graph path
    (64, 0, False, 0.0, 1, 0, (0, 0.0)),        # every step hands its graphs back (default lanes)
    (64, 0, False, 0.0, 1, 1, (0, 0.0)),        # ... to a single lane
    (4, 0, False, 0.02, 200, 0, (0, 0.0)),      # some graphs handed back, staggered
    (8, 0, False, 0.01, 100, 5, (0, 0.0)),      # five lanes, staggered
    (64, 0, False, 0.0, -1, 0, (0, 0.0)),       # no hand-back
    (64, 0, False, 0.0, -1, 0, (-1, -1.0)),     # step-level admission (no slices)
    (5, 0, False, 0.01, -1, 0, (7, -1.0)),      # 7-iteration slices: many resumptions
    (64, 0, False, 0.0, 300, 2, (3, -1.0)),     # 3-iteration slices beside hand-backs
    (64, 0, False, 0.0, -1, 0, (-1, 0.005)),    # 5-microsecond time slices
    (6, "one", False, 0.01, 150, 3, (11, -1.0)),  # pool, slices and hand-backs together
])
def test_serving_same_alone_and_in_mix(gpu_available, tenants, slots, budget, plain, spread,
                                       heavy, lanes, slice):
    check_served(tenants, slots, budget, plain, spread, heavy, lanes, slice)


# after: iteration of a batch from which a sliced graph goes to a hand-back
# lane; heavyFrontier 2**30 keeps the frontier-size hand-back out of it
@pytest.mark.parametrize("after,lanes,slice", [
    (16, 2, (8, -1.0)),    # batches past 16 iterations finish on two lanes
    (1, 0, (1, -1.0)),     # every batch leaves after its first slice (default lanes)
    (-1, 2, (8, -1.0)),    # never: every batch stays in the shared launches
])
def test_serving_long_batches_handed_back(gpu_available, tenants, after, lanes, slice):
    """A batch still running at iteration `after` of its batch leaves the
    shared launch for a hand-back lane at its next slice stop: the same
    labels, iterations and updates as every tenant alone."""
    check_served(tenants, 64, 0, False, 0.0, 1 << 30, lanes, slice, after)


def check_served(tenants, slots, budget, plain, spread, heavy, lanes, slice, after=0):
    ref = alone(tenants, 3000)
    cfgs = [L.make_config(t, maxIterations=3000) for t in tenants]
    rng = np.random.default_rng(11)
    with L.Multi(tenants, cfgs) as m:
        m.set_tuning(heavyFrontier=heavy)
        m.reset()
        m.run_next(list(range(len(tenants))))  # seed batches: tenants become resident
        reqs = requests(tenants, rng, spread)
        if budget == "one":
            budget = max(m.batch_scratch_bytes(t, b) for t, inp in enumerate(tenants)
                         for b in range(1, inp.numBatches))
        print(f"serve slots={slots} budget={budget} plain={plain} heavy={heavy} "
              f"slice={slice} after={after}", flush=True)
        done, steps = m.serve(reqs, maxSlots=slots, hbmBudgetBytes=budget, plain=plain,
                              handBackLanes=lanes, sliceIterations=slice[0], sliceMs=slice[1],
                              handBackAfter=after)
        m.synchronize()
        assert np.all(done >= np.array([r[1] for r in reqs]))
        if plain:
            assert steps == len(reqs)
        if budget:
            st = m.pool_stats()
            assert st["capacity"] == budget // 256 * 256
            assert st["inUse"] == 0 and 0 < st["peak"] <= st["capacity"]
            assert st["leases"] == len(reqs)
            if slots > 1 and spread == 0.0:
                assert st["deferrals"] > 0  # the batches arrived together
        for t, inp in enumerate(tenants):
            assert np.array_equal(bits(m.labels(t)), bits(ref[t][0])), f"tenant {t}"
            for b in range(inp.numBatches):
                s = m.stats(t, b)
                assert (s["iterations"], s["updates"]) == \
                    (ref[t][1][b]["iterations"], ref[t][1][b]["updates"]), f"tenant {t} b {b}"


def test_serving_rejects_bad_config(gpu_available, tenants):
    """Negative limits other than -1 are refused before anything runs."""
    from labelpropcuda_amd import _lib
    with L.Multi(tenants[:2]) as m:
        m.reset()
        m.run_next([0, 1])
        for kw in ({"handBackAfter": -2}, {"sliceIterations": -2}, {"maxSlots": 0}):
            with pytest.raises(_lib.DynLPError):
                m.serve([(0, 0.0)], **kw)
        done, _ = m.serve([(0, 0.0)], handBackAfter=-1)  # still usable
        assert done[0] >= 0.0

