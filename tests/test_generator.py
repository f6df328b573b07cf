"""Synthetic input generator (generateSparseGraph / generateBatches restated
in labelpropcuda_amd/csrc/host/graphgen.cpp): closed-form counts, structural
invariants, and a checksum fixture that catches any drift of the RNG call
sequence."""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import labelpropcuda_amd as L

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "generator_checksums.json")


def digest(inp):
    h = hashlib.sha256()
    for a in (inp.rowPtr, inp.colInd, inp.values, inp.trueLabels, inp.insOff, inp.ins,
              inp.gtOff, inp.gtIds, inp.gtLab, inp.delOff, inp.dels, inp.eOff, inp.eSrc,
              inp.eDst, inp.eW):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("N,deg,B", [(5000, 5, 10), (2000, 3, 4), (1000, 7, 1), (3001, 6, 7)])
def test_graph_invariants(N, deg, B):
    inp = L.generate(N, deg, B, 42)
    rp, ci, w, lab = inp.rowPtr, inp.colInd, inp.values, inp.trueLabels
    assert rp[0] == 0 and rp[-1] == ci.shape[0]
    target = max(N - 1, N * deg // 2)
    assert ci.shape[0] == 2 * target
    rows = np.repeat(np.arange(N), np.diff(rp))
    assert np.all(ci != rows), "self loop"
    for u in range(0, N, max(1, N // 200)):
        r = ci[rp[u]:rp[u + 1]]
        assert np.all(np.diff(r) > 0), "rows sorted, no duplicates"
    A = sp.csr_matrix((w, ci, rp), shape=(N, N))
    assert (A != A.T).nnz == 0, "symmetric with equal weights"
    assert np.sum(lab == 0) == N // 2 and np.sum(lab == 1) == N - N // 2
    same = lab[rows] == lab[ci]
    assert np.all((w[same] >= 0.5) & (w[same] < 1.0))
    assert np.all((w[~same] >= 0.01) & (w[~same] < 0.5))
    assert connected_components(A, directed=False)[0] == 1, "spanning tree"


@pytest.mark.parametrize("N,deg,B", [(5000, 5, 10), (2000, 3, 4), (3001, 6, 7)])
def test_batch_invariants(N, deg, B):
    inp = L.generate(N, deg, B, 42)
    gt = 0.01
    ins_sizes = np.diff(inp.insOff)
    n0 = max(2, int(np.float32(N) * np.float32(gt)))
    assert ins_sizes[0] == 2 * (n0 // 2)
    R = N - ins_sizes[0]
    per, left = divmod(R, max(1, B - 1))
    for b in range(1, B):
        assert ins_sizes[b] == per + (1 if b <= left else 0)
    # every vertex inserted exactly once (partition of V)
    assert np.array_equal(np.sort(inp.ins), np.arange(N))
    batchOf = np.empty(N, np.int64)
    for b in range(B):
        batchOf[inp.ins[inp.insOff[b]:inp.insOff[b + 1]]] = b
    gt_seen = set(inp.gtIds[inp.gtOff[0]:inp.gtOff[1]].tolist())
    assert np.array_equal(np.sort(inp.gtIds[:inp.gtOff[1]]), np.sort(inp.ins[:inp.insOff[1]]))
    active = set()
    deleted = set()
    for b in range(1, B):
        bt = inp.batch(b)
        ngt = max(1, int(np.float32(len(bt["ins"])) * np.float32(gt)))
        assert len(bt["gtIds"]) == min(ngt, len(bt["ins"]))
        assert set(bt["gtIds"].tolist()) <= set(bt["ins"].tolist())
        assert np.array_equal(bt["gtLab"], inp.trueLabels[bt["gtIds"]])
        gt_seen |= set(bt["gtIds"].tolist())
        active |= set(bt["ins"].tolist()) - gt_seen
        nd = min(int(np.float32(len(bt["ins"])) * np.float32(0.09)), len(active))
        assert len(bt["dels"]) == nd
        dset = set(bt["dels"].tolist())
        assert len(dset) == nd and dset <= active and not (dset & gt_seen) and not (dset & deleted)
        active -= dset
        deleted |= dset
    # each full-graph edge appears once, in batch max(b_u, b_v), in CSR order
    rows = np.repeat(np.arange(N), np.diff(inp.rowPtr))
    up = rows < inp.colInd
    exp_b = np.maximum(batchOf[rows[up]], batchOf[inp.colInd[up]])
    for b in range(B):
        bt = inp.batch(b)
        sel = exp_b == b
        assert np.array_equal(bt["eSrc"], rows[up][sel])
        assert np.array_equal(bt["eDst"], inp.colInd[up][sel])
        assert np.array_equal(bt["eW"], inp.values[up][sel])


def test_generator_checksums():
    with open(GOLDEN) as f:
        golden = json.load(f)
    for key, want in golden.items():
        N, deg, B, seed = (int(x) for x in key.split(","))
        assert digest(L.generate(N, deg, B, seed)) == want, key


def test_generator_rejects_bad_arguments():
    with pytest.raises(L.DynLPError):
        L.generate(1, 5, 10, 42)
    with pytest.raises(L.DynLPError):
        L.generate(100, 0, 10, 42)
    with pytest.raises(L.DynLPError):
        L.generate(100, 5, 0, 42)
