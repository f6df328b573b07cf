"""The host-side symmetry test of a full graph (dlp_graph_symmetric): it
decides whether the listed-row CSR update and the multi-wave register-mask
kernels may run, so a false "symmetric" would give wrong results.  Checked
against a direct multiset comparison of (row, column) pairs on random
graphs, with every kind of asymmetry planted once."""
import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import dynlp


def to_csr(n, pairs):
    rows = [[] for _ in range(n)]
    for u, v in pairs:
        rows[u].append(v)
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.array([v for r in rows for v in r], np.int32)
    return rp, ci


def brute(n, rp, ci):
    from collections import Counter
    fwd = Counter()
    for u in range(n):
        for v in ci[rp[u]:rp[u + 1]]:
            if v >= 0:
                fwd[(u, int(v))] += 1
    return all(fwd[(v, u)] == c for (u, v), c in fwd.items())


def random_symmetric(n, m, rng):
    pairs = []
    for u, v in rng.integers(0, n, (m, 2)).tolist():
        if u != v:
            pairs += [(u, v), (v, u)]
    rng.shuffle(pairs)
    return pairs


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_and_planted_asymmetries(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    pairs = random_symmetric(n, int(rng.integers(1, 4 * n)), rng)
    rp, ci = to_csr(n, pairs)
    assert dynlp.graph_symmetric(rp, ci) is True and brute(n, rp, ci)
    if not pairs:
        return
    # one reverse entry dropped
    drop = pairs[: int(rng.integers(0, len(pairs)))] + pairs[int(rng.integers(0, len(pairs))) + 1:]
    rp2, ci2 = to_csr(n, drop)
    assert dynlp.graph_symmetric(rp2, ci2) == brute(n, rp2, ci2)
    # one entry duplicated (multiplicities differ)
    dup = pairs + [pairs[0]]
    rp3, ci3 = to_csr(n, dup)
    assert dynlp.graph_symmetric(rp3, ci3) is False and not brute(n, rp3, ci3)
    # one entry redirected
    u, v = pairs[0]
    w = (v + 1) % n if (v + 1) % n != u else (v + 2) % n
    red = [(u, w)] + pairs[1:]
    rp4, ci4 = to_csr(n, red)
    assert dynlp.graph_symmetric(rp4, ci4) == brute(n, rp4, ci4)


def test_negative_entries_are_ignored_and_self_loops_are_symmetric():
    rp = np.array([0, 3, 5, 6], np.int32)
    ci = np.array([1, -1, 0, 0, -1, 2], np.int32)  # 0-1 both ways, holes, a self loop on 2
    assert dynlp.graph_symmetric(rp, ci) is True
    ci[5] = 1  # 2 -> 1 without 1 -> 2
    assert dynlp.graph_symmetric(rp, ci) is False


def test_edge_cases_and_bad_input():
    assert dynlp.graph_symmetric(np.array([0], np.int32), np.array([], np.int32)) is True
    assert dynlp.graph_symmetric(np.array([0, 0, 0], np.int32), np.array([], np.int32)) is True
    with pytest.raises(L._lib.DynLPError):
        dynlp.graph_symmetric(np.array([0, 2, 1], np.int32), np.array([1, 0], np.int32))


def test_the_generators_graphs_are_symmetric():
    inp = L.generate(3000, 5, 3, 7)
    assert dynlp.graph_symmetric(inp.rowPtr, inp.colInd) is True
