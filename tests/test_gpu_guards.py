"""Regression tests of the counter wrap-around guards and of the lockstep
state check of the multi-graph runner, through the C-ABI: a context whose
iteration stamps and LDS generations start next to their limits must still
match the oracle bit for bit, with both counters rebased on the way."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import _lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1
GEN_LIMIT = 2**31 - 2


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def near_limits(max_iter):
    """stampBase and localGen from which the first batch fits and the second
    needs a rebase and a restart."""
    return INT_MAX - 1 - (max_iter + 1) - 50, GEN_LIMIT - (4 * (max_iter + 1) + 4) - 100


def test_context_guards_near_limits_match_oracle(gpu_available, c1_inputs):
    max_iter = 2000
    inp = c1_inputs
    with L.Context(inp, L.make_config(inp, maxIterations=max_iter)) as ctx:
        ctx.reset()
        sb0, lg0 = near_limits(max_iter)
        ctx.set_counter_bases(sb0, lg0)
        assert ctx.counter_bases() == (sb0, lg0)
        orc = Oracle(inp, maxIterations=max_iter)
        for b in range(inp.numBatches):
            ctx.run_batch(b)
            orc.run_batch(b)
            assert np.array_equal(bits(ctx.labels()), bits(orc.labels())), f"batch {b}"
            st, ost = ctx.stats(b), orc.stats()
            assert (st["iterations"], st["updates"]) == (ost["iterations"], ost["updates"])
        sb, lg = ctx.counter_bases()
        assert sb < 10 * (max_iter + 2), "stamps were not rebased"
        assert lg < lg0, "generations were not restarted"


def test_multi_guards_near_limits_match_plain(gpu_available):
    graphs = [L.generate(3000, 5, 5, 31), L.generate(5000, 3, 5, 32)]
    max_iter = 1500
    cfgs = [L.make_config(g, maxIterations=max_iter) for g in graphs]
    plain = []
    for g, c in zip(graphs, cfgs):
        with L.Context(g, c) as ctx:
            ctx.run_all()
            plain.append(ctx.labels())
    for mode in (1, 2):  # captured graphs, enqueue every time
        with L.Multi(graphs, cfgs) as m:
            m.set_tuning(useGraphs=mode)
            m.reset()
            sb0, lg0 = near_limits(max_iter)
            m.set_counter_bases(1, sb0, lg0)
            for b in range(5):
                m.run_batch(b)
            m.synchronize()
            for g in range(2):
                assert np.array_equal(bits(m.labels(g)), bits(plain[g])), f"graph {g} mode {mode}"
            sb, lg = m.counter_bases(1)
            assert sb < 10 * (max_iter + 2) and lg < lg0


def test_multi_run_batch_refused_after_run_next(gpu_available):
    graphs = [L.generate(2000, 5, 4, 41), L.generate(2500, 5, 4, 42)]
    with L.Multi(graphs) as m:
        m.reset()
        m.run_next([0, 1])
        with pytest.raises(L.DynLPError) as e:
            m.run_batch(0)
        assert e.value.code == _lib.DLP_ERR_STATE
        with pytest.raises(L.DynLPError) as e:
            m.run_batch(1)
        assert e.value.code == _lib.DLP_ERR_STATE
        m.reset()  # lockstep again after a reset
        m.run_batch(0)
        m.synchronize()
