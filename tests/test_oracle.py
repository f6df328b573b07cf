"""Pins for the CPU oracle (oracle/dynlp_oracle.c).

Each check below can fail for a plausible oracle mistake: the propagation
against an independent thread-level restatement of the reference kernels
(tests/refsim.py) on random graphs, rows longer than 256 and real batch
states; CC against scipy; sparsification as a set identity; the CSR update
against a brute-force rebuild; label initialisation and propagation against
closed forms; the double-buffer staleness of the reference as a hand-worked
case; the FP64 twin as an error bound.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components as scipy_cc

import oracle.oracle as O
from refsim import ref_active_csr, ref_mark_deletions, ref_propagate

F32 = np.float32


def random_graph(rng, n, p, hub=0):
    """Symmetric CSR with sorted rows; optional hub vertex 0 adjacent to `hub` others."""
    A = np.triu(rng.random((n, n)) < p, 1)
    if hub:
        A[0, 1:hub + 1] = True
    A = A | A.T
    rows, cols = np.nonzero(A)
    w = rng.uniform(0.01, 1.0, size=(n, n)).astype(F32)
    w = np.triu(w, 1)
    w = w + w.T
    rowPtr = np.zeros(n + 1, np.int32)
    np.add.at(rowPtr, rows + 1, 1)
    rowPtr = np.cumsum(rowPtr).astype(np.int32)
    return rowPtr, cols.astype(np.int32), w[rows, cols].astype(F32)


def random_state(rng, n, colInd, hole_frac=0.1):
    status = rng.choice([0, 0, 0, 0, 1, 2], size=n).astype(np.int32)
    gtClass = np.where(status == 2, rng.integers(0, 2, n), -1).astype(np.int32)
    labels = rng.random(n).astype(F32)
    labels[gtClass == 0] = 0
    labels[gtClass == 1] = 1
    aff = (rng.random(n) < 0.5).astype(np.int32)
    col = colInd.copy()
    col[rng.random(col.shape[0]) < hole_frac] = -1
    return status, gtClass, labels, aff, col


def same_bits(a, b):
    return a.dtype == b.dtype and np.array_equal(a.view(np.int32 if a.dtype == F32 else np.int64),
                                                 b.view(np.int32 if b.dtype == F32 else np.int64))


@pytest.mark.parametrize("seed", range(6))
def test_propagate_matches_kernel_restatement(seed):
    rng = np.random.default_rng(seed)
    n = 60
    rowPtr, colInd, values = random_graph(rng, n, 0.12)
    status, gtClass, labels, aff, col = random_state(rng, n, colInd)
    delta = float(rng.choice([1e-4, 1e-3, 1e-2]))
    for cap in (1, 2, 7, 40):
        got = O.propagate(rowPtr, col, values, labels, status, gtClass, aff, delta, cap)
        exp = ref_propagate(rowPtr, col, values, labels, status, gtClass, aff, delta, cap)
        assert got[3] == exp[3], "iterationsUsed"
        assert np.array_equal(got[4], exp[4]), "frontier sizes"
        assert same_bits(got[0], exp[0]), "labels"
        assert same_bits(got[1], exp[1]), "labelsBuffer"
        assert np.array_equal(got[2], exp[2]), "isAffected exit state"


def test_propagate_rows_longer_than_256_slots():
    # vertex 0 has 300 neighbours: slots 0..43 hold two positions each
    rng = np.random.default_rng(7)
    n = 320
    rowPtr, colInd, values = random_graph(rng, n, 0.01, hub=300)
    assert rowPtr[1] - rowPtr[0] > 256
    status = np.zeros(n, np.int32)
    gtClass = np.full(n, -1, np.int32)
    gt = rng.choice(np.arange(1, n), 40, replace=False)
    status[gt] = 2
    gtClass[gt] = rng.integers(0, 2, 40)
    labels = rng.random(n).astype(F32)
    labels[gtClass == 0] = 0
    labels[gtClass == 1] = 1
    aff = np.ones(n, np.int32)
    got = O.propagate(rowPtr, colInd, values, labels, status, gtClass, aff, 1e-4, 25)
    exp = ref_propagate(rowPtr, colInd, values, labels, status, gtClass, aff, 1e-4, 25)
    assert got[3] == exp[3] and np.array_equal(got[4], exp[4])
    assert same_bits(got[0], exp[0]) and same_bits(got[1], exp[1])


def test_propagate_matches_restatement_on_c1_batch_state(c1_inputs):
    o = O.Oracle(c1_inputs)
    for b in range(4):
        o.run_batch(b)
    rp, ci, vv = o.csr()
    init, aff = o.init_labels(), o.affected()
    d, _, g = o.status()
    got = O.propagate(rp, ci, vv, init, d, g, aff, 1e-4, 120)
    exp = ref_propagate(rp, ci, vv, init, d, g, aff, 1e-4, 120)
    assert got[3] == exp[3] == 120
    assert np.array_equal(got[4], exp[4])
    assert same_bits(got[0], exp[0]) and same_bits(got[1], exp[1])
    assert np.array_equal(got[2], exp[2])


def test_propagate_closed_form_two_gt_neighbours():
    # u=0 with GT0 neighbour (w=a) and GT1 neighbour (w=b): one update to
    # b/(a+b), then no change -> 2 iterations, label b/(a+b)
    a, b = F32(0.3), F32(0.7)
    rowPtr = np.array([0, 2, 3, 4], np.int32)
    colInd = np.array([1, 2, 0, 0], np.int32)
    values = np.array([a, b, a, b], F32)
    status = np.array([0, 2, 2], np.int32)
    gtClass = np.array([-1, 0, 1], np.int32)
    labels = np.array([0.5, 0, 1], F32)
    aff = np.array([1, 0, 0], np.int32)
    lab, _, _, it, sizes = O.propagate(rowPtr, colInd, values, labels, status, gtClass, aff,
                                       1e-4, 100)
    assert it == 2 and sizes.tolist() == [1, 1]
    assert lab[0] == F32(b) / F32(a + b)
    assert lab[1] == 0 and lab[2] == 1


def test_propagate_isolated_vertex_keeps_label():
    rowPtr = np.array([0, 0, 1, 2], np.int32)
    colInd = np.array([2, 1], np.int32)
    values = np.array([0.5, 0.5], F32)
    status = np.zeros(3, np.int32)
    gtClass = np.full(3, -1, np.int32)
    labels = np.array([0.25, 0.4, 0.6], F32)
    lab, _, _, it, _ = O.propagate(rowPtr, colInd, values, labels, status, gtClass,
                                   np.array([1, 0, 0], np.int32), 1e-4, 100)
    assert it == 1 and lab[0] == F32(0.25)


def test_propagate_double_buffer_revert():
    """The reference's stale write buffer: a vertex whose update is <= delta
    leaves the frontier, and when the loop ends on the other buffer its
    update is lost (iterativePropagation.cu:308-334).
      u (=0): one GT1 neighbour, label 1 - 4e-5 -> update to 1.0, |d| <= delta
      w (=2): one GT0 neighbour, label 0.5 -> 0.0 (changed), then stable
    iteration 0: {u, w}; iteration 1: {w}; last written buffer = labels ->
    u keeps its old value."""
    rowPtr = np.array([0, 1, 2, 3, 4], np.int32)
    colInd = np.array([1, 0, 3, 2], np.int32)
    values = np.array([1.0, 1.0, 1.0, 1.0], F32)
    status = np.array([0, 2, 0, 2], np.int32)
    gtClass = np.array([-1, 1, -1, 0], np.int32)
    u0 = F32(1.0) - F32(4e-5)
    labels = np.array([u0, 1, 0.5, 0], F32)
    lab, buf, _, it, sizes = O.propagate(rowPtr, colInd, values, labels, status, gtClass,
                                         np.array([1, 0, 1, 0], np.int32), 1e-4, 100)
    assert it == 2 and sizes.tolist() == [2, 1]
    assert lab[0] == u0          # reverted
    assert buf[0] == F32(1.0)    # the update survives only in the other buffer
    assert lab[2] == 0


def test_propagate_fixed_point_invariants(c1_inputs):
    o = O.Oracle(c1_inputs, maxIterations=100000)
    for b in range(3):
        o.run_batch(b)
    lab = o.labels()
    _, hs, g = o.status()
    assert np.all((lab >= 0) & (lab <= 1))
    assert np.all(lab[g == 0] == 0) and np.all(lab[g == 1] == 1)


def test_mark_deletions_matches_restatement():
    rng = np.random.default_rng(3)
    for trial in range(5):
        n = 50
        rowPtr, colInd, _ = random_graph(rng, n, 0.15)
        status = rng.choice([0, 0, 0, 2], size=n).astype(np.int32)
        aff = (rng.random(n) < 0.2).astype(np.int32) * rng.integers(1, 3, n).astype(np.int32)
        deleted = rng.choice(n, 6, replace=False).astype(np.int32)
        deleted = np.concatenate([deleted, deleted[:2], [-3, n + 5]]).astype(np.int32)
        got = O.mark_deletions(rowPtr, colInd, status, aff, deleted, n)
        exp = ref_mark_deletions(rowPtr, colInd, status, aff, deleted, n)
        for x, y in zip(got, exp):
            assert np.array_equal(x, y)


def test_sparsify_set_identity():
    rng = np.random.default_rng(5)
    n, m = 400, 1500
    src = rng.integers(0, n, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    w = rng.random(m).astype(F32)
    isNew = (rng.random(n) < 0.4).astype(np.int32)
    numInserted = 150
    tau = F32(0.5)
    ss, sd, sw, aff = O.sparsify(src, dst, w, numInserted, isNew, tau, np.zeros(n, np.int32))
    keep = (w > tau) & (isNew[src] == 1) & (isNew[dst] == 1)
    exp_src = np.stack([src[keep], dst[keep]], 1).reshape(-1)
    exp_dst = np.stack([dst[keep], src[keep]], 1).reshape(-1)
    exp_w = np.repeat(w[keep], 2)
    assert np.array_equal(ss, exp_src) and np.array_equal(sd, exp_dst)
    assert np.array_equal(sw, exp_w)
    heavy = w > tau
    marked = np.zeros(n, np.int32)
    for x in np.concatenate([src[heavy], dst[heavy]]):
        if x < numInserted:
            marked[x] = 1
    assert np.array_equal(aff, marked)


@pytest.mark.parametrize("seed", range(4))
def test_connected_components_match_scipy(seed):
    rng = np.random.default_rng(seed)
    N = 3000
    inserted = rng.choice(N, 800, replace=False).astype(np.int32)
    k = 500
    a = rng.choice(inserted, k)
    b = rng.choice(inserted, k)
    src = np.concatenate([a, b]).astype(np.int32)
    dst = np.concatenate([b, a]).astype(np.int32)
    comp, nc = O.connected_components(src, dst, inserted, N)
    # scipy on the inserted vertices
    idx = {v: i for i, v in enumerate(inserted.tolist())}
    A = sp.coo_matrix((np.ones(2 * k), ([idx[x] for x in src], [idx[x] for x in dst])),
                      shape=(800, 800))
    ncs, lab = scipy_cc(A, directed=False)
    assert nc == ncs
    # canonical ids: rank of each component's minimum vertex id
    mins = {}
    for v, l in zip(inserted.tolist(), lab.tolist()):
        mins[l] = min(mins.get(l, v), v)
    order = {m: r for r, m in enumerate(sorted(mins.values()))}
    exp = np.full(N, -1, np.int32)
    for v, l in zip(inserted.tolist(), lab.tolist()):
        exp[v] = order[mins[l]]
    assert np.array_equal(comp, exp)


def test_connected_components_edge_cases():
    comp, nc = O.connected_components(np.zeros(0, np.int32), np.zeros(0, np.int32),
                                      np.array([3, 1], np.int32), 5)
    assert nc == 2 and np.all(comp == -1)
    comp, nc = O.connected_components(np.zeros(0, np.int32), np.zeros(0, np.int32),
                                      np.zeros(0, np.int32), 5)
    assert nc == 0
    with pytest.raises(ValueError):
        O.connected_components(np.array([0], np.int32), np.array([4], np.int32),
                               np.array([0, 1], np.int32), 5)


def test_label_initialization_closed_forms():
    # vertices 0..5; GT: 4 (class 0), 5 (class 1).  Components: {0,1,2} and {3}.
    # rows: 0:[4(w=.2),5(w=.7)], 1:[5(.3), -1], 2:[4(.9)], 3:[]
    rowPtr = np.array([0, 2, 4, 5, 5, 5, 5], np.int32)
    colInd = np.array([4, 5, 5, -1, 4], np.int32)
    values = np.array([0.2, 0.7, 0.3, 0.8, 0.9], F32)
    comp = np.array([0, 0, 0, 1, -1, -1], np.int32)
    gt = np.array([-1, -1, -1, -1, 0, 1], np.int32)
    ins = np.array([2, 0, 3, 1], np.int32)
    lab = O.label_initialization(rowPtr, colInd, values, np.full(6, 0.5, F32), comp, 2, ins, gt)
    # member sums in ascending id: p0 = (.2, 0, .9), p1 = (.7, .3, 0); tree ((p_0+p_1)+p_2)
    W0 = (F32(0.2) + F32(0)) + F32(0.9)
    W1 = (F32(0.7) + F32(0.3)) + F32(0)
    f = W1 / (W0 + W1)
    assert lab[0] == f and lab[1] == f and lab[2] == f
    assert lab[3] == F32(0.5)            # no GT neighbour
    # GT members are set to 0/1; a vertex without a component gets 0.5
    gt2 = gt.copy()
    gt2[1] = 1
    comp2 = comp.copy()
    comp2[3] = -1
    lab2 = O.label_initialization(rowPtr, colInd, values, np.full(6, 0.5, F32), comp2, 2, ins,
                                  gt2)
    assert lab2[1] == 1 and lab2[3] == F32(0.5)


def test_label_initialization_fp64_agreement(c1_inputs):
    o32 = O.Oracle(c1_inputs)
    o64 = O.Oracle(c1_inputs, fp64=True)
    for b in range(3):
        o32.run_batch(b)
        o64.run_batch(b)
        i32, i64 = o32.init_labels(), o64.init_labels()
        assert np.max(np.abs(i32.astype(np.float64) - i64)) < 1e-6


def test_active_csr_matches_bruteforce_rebuild(c1_inputs):
    inp = c1_inputs
    N = inp.numNodes
    batchMap = np.full(N, -1, np.int32)
    for b in range(inp.numBatches):
        batchMap[inp.ins[inp.insOff[b]:inp.insOff[b + 1]]] = b
    o = O.Oracle(inp)
    for b in range(inp.numBatches):
        o.run_batch(b)
        d, hs, g = o.status()
        rp, ci, vv = ref_active_csr(inp.rowPtr, inp.colInd, inp.values, hs, batchMap, b)
        bt = inp.batch(b)
        # device status before markDeletions is the host status
        ci2, st2, aff = ref_mark_deletions(rp, ci, hs, np.zeros(N, np.int32), bt["dels"], N)
        orp, oci, ovv = o.csr()
        assert np.array_equal(orp, rp) and np.array_equal(oci, ci2) and np.array_equal(ovv, vv)
        assert np.array_equal(st2, d)


def test_fp64_twin_error_is_small(c1_inputs):
    o32 = O.Oracle(c1_inputs, maxIterations=2000)
    o64 = O.Oracle(c1_inputs, maxIterations=2000, fp64=True)
    o32.run_all()
    o64.run_all()
    err = np.abs(o32.labels().astype(np.float64) - o64.labels())
    # threshold decisions may differ, so this is a statistic, bounded loosely
    assert err.mean() < 1e-3 and err.max() < 0.1


@pytest.mark.parametrize("fp64", [False, True])
def test_parallel_propagation_gives_identical_results(c1_inputs, fp64):
    """The frontier-parallel variant (the multi-core CPU baseline) against
    the sequential restatement: same labels, frontier sizes, iteration and
    update counts, accuracy."""
    def run(threads):
        O.set_threads(threads, minFrontier=1)
        try:
            o = O.Oracle(c1_inputs, fp64=fp64)
            st = []
            for b in range(c1_inputs.numBatches):
                o.run_batch(b)
                st.append((o.stats(), o.frontier_sizes().tolist(), o.affected().tolist()))
            return o.labels(), st
        finally:
            O.set_threads(1)
    lab1, st1 = run(1)
    lab4, st4 = run(4)
    assert same_bits(lab1, lab4)
    assert st1 == st4
