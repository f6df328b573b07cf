"""Pins for the oracle's IrLP (oracle/dynlp_oracle.c irlp, oracle.py
irlp_final_state).

- bit-exact against the thread-level restatement of irLPKernel /
  irLPEnforceGTKernel / launchIrLP (tests/refsim.py ref_irlp) on random
  graphs with holes, deleted neighbours, ground truth and rows longer than
  256 (a dropped status check, a wrong tree order or an in-place update
  fails here);
- the converged labels against the harmonic solution of the graph
  Laplacian (numpy.linalg.solve), in FP64, and components without ground
  truth staying at exactly 0.5 (a wrong weight or a missing GT enforcement
  fails here);
- the loop's iteration count and final-buffer choice against hand-worked
  cases (a dropped swap or an off-by-one in iterationsUsed fails here);
- the final-state replay against a hand-built batch stream with the
  reference's insert -> delete -> ground-truth order (main.cu:131-146).
"""
from types import SimpleNamespace

import numpy as np
import pytest

import oracle.oracle as O
from refsim import ref_irlp
from test_oracle import F32, random_graph, random_state, same_bits


@pytest.mark.parametrize("seed", range(6))
def test_irlp_matches_kernel_restatement(seed):
    rng = np.random.default_rng(100 + seed)
    n = 70
    rowPtr, colInd, values = random_graph(rng, n, 0.1)
    status, gtClass, labels, _, col = random_state(rng, n, colInd)
    for delta, maxIt in ((1e-4, 200), (1e-2, 200), (0.0, 17)):
        got, gi = O.irlp(rowPtr, col, values, labels, status, gtClass, delta, maxIt)
        exp, ei = ref_irlp(rowPtr, col, values, labels, status, gtClass, delta, maxIt)
        assert gi == ei
        assert same_bits(got, exp)


def test_irlp_rows_longer_than_256_slots():
    rng = np.random.default_rng(11)
    n = 330
    rowPtr, colInd, values = random_graph(rng, n, 0.01, hub=310)
    assert rowPtr[1] - rowPtr[0] > 256
    status = np.zeros(n, np.int32)
    gtClass = np.full(n, -1, np.int32)
    gt = rng.choice(np.arange(1, n), 30, replace=False)
    status[gt] = 2
    gtClass[gt] = rng.integers(0, 2, 30)
    labels = np.full(n, 0.5, F32)
    labels[gtClass == 0] = 0
    labels[gtClass == 1] = 1
    got, gi = O.irlp(rowPtr, colInd, values, labels, status, gtClass, 1e-5, 300)
    exp, ei = ref_irlp(rowPtr, colInd, values, labels, status, gtClass, 1e-5, 300)
    assert gi == ei > 2
    assert same_bits(got, exp)


def harmonic(rowPtr, colInd, values, gtClass):
    """Labels of the unlabelled vertices solving L_uu f_u = -L_ul f_l."""
    n = rowPtr.shape[0] - 1
    A = np.zeros((n, n))
    for u in range(n):
        for e in range(rowPtr[u], rowPtr[u + 1]):
            A[u, colInd[e]] += values[e]
    lab = gtClass >= 0
    f = np.where(gtClass == 1, 1.0, 0.0)
    U = np.nonzero(~lab)[0]
    L = np.diag(A.sum(1)) - A
    f[U] = np.linalg.solve(L[np.ix_(U, U)], -L[np.ix_(U, np.nonzero(lab)[0])] @ f[lab])
    return f


def test_irlp_converges_to_harmonic_solution():
    rng = np.random.default_rng(5)
    n = 40
    rowPtr, colInd, values = random_graph(rng, n, 0.2)
    # connected: add a path through all vertices
    import scipy.sparse as sp
    A = sp.csr_matrix((values, colInd, rowPtr), shape=(n, n)).toarray()
    for i in range(n - 1):
        if A[i, i + 1] == 0:
            A[i, i + 1] = A[i + 1, i] = 0.5
    M = sp.csr_matrix(A)
    rowPtr, colInd, values = (M.indptr.astype(np.int32), M.indices.astype(np.int32),
                              M.data.astype(F32))
    gtClass = np.full(n, -1, np.int32)
    gtClass[[0, 7]] = 0
    gtClass[[20, 33]] = 1
    status = np.where(gtClass >= 0, 2, 0).astype(np.int32)
    labels = np.where(gtClass >= 0, gtClass, 0.5).astype(F32)
    got, it = O.irlp(rowPtr, colInd, values, labels, status, gtClass, 1e-13, 200000, fp64=True)
    assert it < 200000
    exp = harmonic(rowPtr, colInd, values.astype(np.float64), gtClass)
    assert np.max(np.abs(got - exp)) < 1e-9


def test_irlp_path_is_linear_and_unlabelled_component_stays_half():
    # path 0-1-...-9 with GT 0 at 0 and GT 1 at 9 (unit weights): f_i = i/9;
    # a separate triangle 10-11-12 with no GT keeps 0.5 exactly
    edges = [(i, i + 1) for i in range(9)] + [(10, 11), (11, 12), (10, 12)]
    n = 13
    rows = [[] for _ in range(n)]
    for a, b in edges:
        rows[a].append(b)
        rows[b].append(a)
    rowPtr = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    colInd = np.concatenate([sorted(r) for r in rows]).astype(np.int32)
    values = np.ones(colInd.shape[0], F32)
    gtClass = np.full(n, -1, np.int32)
    gtClass[0], gtClass[9] = 0, 1
    status = np.where(gtClass >= 0, 2, 0).astype(np.int32)
    labels = np.full(n, 0.5, F32)
    labels[0], labels[9] = 0, 1
    got, _ = O.irlp(rowPtr, colInd, values, labels, status, gtClass, 1e-14, 100000, fp64=True)
    assert np.max(np.abs(got[:10] - np.arange(10) / 9.0)) < 1e-11
    got32, _ = O.irlp(rowPtr, colInd, values, labels, status, gtClass, 1e-4, 100000)
    assert np.all(got32[10:] == F32(0.5))


def tiny_chain():
    # 0 - 1 - 2 with GT1 at 0: iteration 1 sets 1 -> (1+0.5)/2 = 0.75,
    # 2 -> 0.5; iteration 2: 1 -> (1+0.5)/2 = 0.75, 2 -> 0.75; ...
    rowPtr = np.array([0, 1, 3, 4], np.int32)
    colInd = np.array([1, 0, 2, 1], np.int32)
    values = np.ones(4, F32)
    status = np.array([2, 0, 0], np.int32)
    gtClass = np.array([1, -1, -1], np.int32)
    labels = np.array([1.0, 0.5, 0.5], F32)
    return rowPtr, colInd, values, labels, status, gtClass


def test_irlp_iteration_count_and_final_buffer():
    rp, ci, vv, lab, st, gt = tiny_chain()
    out, it = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 0)
    assert it == 0 and np.array_equal(out, lab)
    out, it = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 1)  # odd count: result of iteration 0
    assert it == 1 and np.array_equal(out, np.array([1.0, 0.75, 0.5], F32))
    out, it = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 2)  # even count: back in labels
    assert it == 2 and np.array_equal(out, np.array([1.0, 0.75, 0.75], F32))
    # a delta above every move stops after the first iteration, which counts
    out, it = O.irlp(rp, ci, vv, lab, st, gt, 0.3, 100)
    assert it == 1 and np.array_equal(out, np.array([1.0, 0.75, 0.5], F32))
    # converges to all ones; the last iteration is the first with no move > delta
    out, it = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 1000)
    assert np.all(np.abs(out - 1.0) < 1e-3) and 10 < it < 1000


def test_irlp_keeps_deleted_and_enforces_gt():
    rp, ci, vv, lab, st, gt = tiny_chain()
    st = st.copy()
    st[2] = 1                     # vertex 2 DELETED: keeps 0.5, ignored as neighbour
    lab = lab.copy()
    lab[0] = 0.3                  # GT class 1 with a stale label: forced to 1
    out, it = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 100)
    assert out[0] == 1.0 and out[2] == F32(0.5) and out[1] == 1.0
    # iteration 0 reads the stale 0.3 (the GT copy happens before enforcement)
    one, _ = O.irlp(rp, ci, vv, lab, st, gt, 1e-4, 1)
    assert one[1] == F32(0.3) and one[0] == 1.0


def test_final_state_replay_order():
    """Insert -> ACTIVE, delete -> DELETED, GT -> GROUND_TRUTH per batch, in
    that order; a vertex inserted and deleted in one batch ends DELETED; a
    GT vertex deleted later keeps its class but leaves the graph."""
    # full graph: path 0-1-2-3-4-5
    rowPtr = np.array([0, 1, 3, 5, 7, 9, 10], np.int32)
    colInd = np.array([1, 0, 2, 1, 3, 2, 4, 3, 5, 4], np.int32)
    values = np.arange(1, 11, dtype=F32) / 10
    batches = [
        {"ins": np.array([0, 1, 2], np.int32), "dels": np.array([], np.int32),
         "gtIds": np.array([0], np.int32), "gtLab": np.array([0.9], F32)},
        {"ins": np.array([3, 4], np.int32), "dels": np.array([4], np.int32),
         "gtIds": np.array([2], np.int32), "gtLab": np.array([0.2], F32)},
        {"ins": np.array([], np.int32), "dels": np.array([0], np.int32),
         "gtIds": np.array([], np.int32), "gtLab": np.array([], F32)},
    ]
    inputs = SimpleNamespace(numNodes=6, numBatches=3, rowPtr=rowPtr, colInd=colInd,
                             values=values, batch=lambda b: batches[b])
    rp, ci, vv, st, gt, lab = O.irlp_final_state(inputs)
    assert st.tolist() == [1, 0, 2, 0, 1, 1]           # 5 never inserted: DELETED
    assert gt.tolist() == [1, -1, 0, -1, -1, -1]
    assert lab.tolist() == [np.float32(0.9), 0.5, np.float32(0.2), 0.5, 0.5, 0.5]
    # rows of 1, 2, 3 only, entries between them
    assert rp.tolist() == [0, 0, 1, 3, 4, 4, 4]
    assert ci.tolist() == [2, 1, 3, 2]
    assert np.array_equal(vv, values[[2, 3, 4, 5]])
