"""The serving workload at its configured size: 64 device-resident tenant
graphs, vertex counts log-uniform in [10,000, 5,000,000], average degree 5,
128 update batches of log-uniform 100 to 100,000 vertices (90/1/9), tenant t
generated with seed 1042 + t (the benchmark's workload).  Every tenant's
labels and per-batch iteration / update counts must be the same served
alone (one batch at a time through the single-graph path) and in the mix
(coalesced steps with hand-back lanes), for saturated and staggered
arrivals; the tenants small enough for the CPU are also checked against the
oracle."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
T, R = 64, 128


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def c5():
    rng = np.random.default_rng(1042)
    sizesN = np.rint(np.exp(rng.uniform(np.log(1e4), np.log(5e6), T)))
    reqTenant = rng.integers(0, T, R)
    reqSize = np.rint(np.exp(rng.uniform(np.log(100), np.log(1e5), R))).astype(int)
    tenants = [L.generate_stream(int(sizesN[t]), 5,
                                 [0] + [int(k) for k in reqSize[reqTenant == t]], 1042 + t)
               for t in range(T)]
    assert min(g.numNodes for g in tenants) >= 10000
    assert max(g.numNodes for g in tenants) <= 5000000
    return tenants, reqTenant


def served(m, tenants, reqTenant, arrivals, plain, slots=64):
    m.reset()
    m.run_next(list(range(T)))  # seed batches: the tenants become resident
    m.synchronize()
    reqs = list(zip(reqTenant.tolist(), arrivals.tolist()))
    done, steps = m.serve(reqs, maxSlots=slots, plain=plain)
    m.synchronize()
    assert np.all(done >= arrivals)
    return [(m.labels(t), [(s["iterations"], s["updates"])
                           for s in (m.stats(t, b) for b in range(tenants[t].numBatches))])
            for t in range(T)], steps


@pytest.mark.timeout(900)
def test_c5_each_tenant_same_alone_and_in_mix(gpu_available, c5):
    tenants, reqTenant = c5
    with L.Multi(tenants) as m:
        alone, steps = served(m, tenants, reqTenant, np.zeros(R), plain=True)
        assert steps == R
        mixes = {"saturated": np.zeros(R)}
        # staggered: Poisson arrivals at about twice the plain path's rate
        rng = np.random.default_rng(5)
        gaps = rng.exponential(0.004, R)
        mixes["staggered"] = np.cumsum(gaps) - gaps[0]
        for name, arrivals in mixes.items():
            mix, steps = served(m, tenants, reqTenant, arrivals, plain=False)
            assert steps < R or name == "staggered"
            for t in range(T):
                assert np.array_equal(bits(mix[t][0]), bits(alone[t][0])), f"{name}: tenant {t}"
                assert mix[t][1] == alone[t][1], f"{name}: tenant {t} per-batch counts"
    # the oracle on the tenants of at most 100,000 vertices
    small = [t for t in range(T) if tenants[t].numNodes <= 100000]
    assert len(small) >= 8
    for t in small:
        orc = Oracle(tenants[t])
        for b in range(tenants[t].numBatches):
            orc.run_batch(b)
            assert (orc.stats()["iterations"], orc.stats()["updates"]) == alone[t][1][b], \
                f"tenant {t} batch {b}"
        assert np.array_equal(bits(orc.labels()), bits(alone[t][0])), f"tenant {t}: oracle"
