"""The wrap-around guards of the two device counters that only grow across
batches (iteration stamps, LDS working-set generations), as host arithmetic:
no batch may start with a bound from which it could reach the limit, and a
rebase happens only when needed.  No GPU."""
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import dynlp

INT_MAX = 2**31 - 1
GEN_LIMIT = 2**31 - 2


def gens_per_batch(max_iter):
    return 4 * (max_iter + 1) + 4


@pytest.mark.parametrize("max_iter", [0, 1, 100, 100000, 2**28])
def test_guard_never_lets_a_batch_cross_the_limits(max_iter):
    sb, gb = 1, 0
    rebases = restarts = 0
    per_stamp = max_iter + 1
    # enough batches to cross each limit at least twice
    n = min(3 * (INT_MAX // per_stamp) + 5, 200000)
    for _ in range(n):
        p = dynlp.counter_guard_plan(sb, gb, max_iter)
        start_s = 1 if p["rebaseStamps"] else sb
        start_g = 0 if p["restartGens"] else gb
        # every stamp the batch can use stays below the ineligible marker
        assert start_s + per_stamp <= INT_MAX - 1
        assert start_g + gens_per_batch(max_iter) < GEN_LIMIT
        # rebase only when the un-rebased batch could have crossed
        assert p["rebaseStamps"] == (sb + per_stamp >= INT_MAX - 1)
        assert p["restartGens"] == (gb + gens_per_batch(max_iter) >= GEN_LIMIT)
        assert p["stampBound"] == start_s + per_stamp
        assert p["genBound"] == start_g + gens_per_batch(max_iter)
        rebases += p["rebaseStamps"]
        restarts += p["restartGens"]
        sb, gb = p["stampBound"], p["genBound"]
    if max_iter >= 100000:
        assert rebases >= 2 and restarts >= 2


def test_guard_at_the_edge():
    # a batch of 100,000 iterations from stampBase INT_MAX - 100,003 still fits
    p = dynlp.counter_guard_plan(INT_MAX - 100003, 0, 100000)
    assert not p["rebaseStamps"] and p["stampBound"] == INT_MAX - 2
    p = dynlp.counter_guard_plan(INT_MAX - 100002, 0, 100000)
    assert p["rebaseStamps"] and p["stampBound"] == 100002


def test_guard_rejects_unbounded_iteration_caps():
    with pytest.raises(L.DynLPError):
        dynlp.counter_guard_plan(1, 0, 2**28 + 1)


def test_context_rejects_iteration_cap_above_limit(c1_inputs):
    cfg = L.make_config(c1_inputs, maxIterations=2**28 + 1)
    with pytest.raises(L.DynLPError) as e:
        L.Context(c1_inputs, cfg)
    assert "maxIterations" in str(e.value)
