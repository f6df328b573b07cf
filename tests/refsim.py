"""Independent numpy restatements of the reference kernels, for pinning the
oracle on small inputs.

These follow the CUDA sources thread by thread rather than the oracle's C
(which shortens the 256-slot tree to pow2ceil(rowLength) slots, restates SV
by union-find, and so on), so agreement is evidence that the oracle's
restructurings are exact:

  ref_propagate      <- iterativePropagation.cu:241-346 (+ kernels :63-235):
                        a 256-thread block per frontier vertex, slot t summing
                        row positions t, t+256, ... then the shared-memory
                        tree s[t] += s[t+stride], stride = 128 .. 1, in
                        float32; pointer-swapped buffers; final copy; GT
                        enforcement.
  ref_mark_deletions <- sparsifySubgraph.cu:58-111, :185-203
  ref_active_csr     <- dynLP.cu:111-167 (updateHostCSR)
  ref_irlp           <- validation.cu:47-205: the same block row sums over
                        every active non-ground-truth vertex per iteration,
                        then ground-truth enforcement, pointer swap and a
                        changed flag.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def ref_propagate(rowPtr, colInd, values, labels, status, gtClass, isAffected, delta,
                  maxIterations):
    """Returns (labels, labelsBuffer, isAffected, iterationsUsed, frontierSizes)."""
    rowPtr = np.asarray(rowPtr, np.int64)
    colInd = np.asarray(colInd, np.int64)
    values = np.asarray(values, F32)
    status = np.asarray(status)
    gtClass = np.asarray(gtClass)
    N = rowPtr.shape[0] - 1
    bufA = np.array(labels, F32)           # d_labels
    bufB = bufA.copy()                      # d_labelsBuffer (copied at :268-270)
    old, new = bufA, bufB
    aff = np.array(isAffected, np.int32)
    d = F32(delta)
    sizes = []
    iterationsUsed = 0
    for it in range(maxIterations):
        # collectAffectedKernel (:63-79)
        lst = np.nonzero((aff == 1) & (status == 0) & (gtClass < 0))[0]
        if lst.size == 0:
            break
        sizes.append(int(lst.size))
        # propagateLabelsKernel (:98-179): slot array per block
        W, S = block_row_sums(rowPtr, colInd, values, status, old, lst)
        oldLab = old[lst]
        with np.errstate(divide="ignore", invalid="ignore"):
            q = (S / W).astype(F32)
        nl = np.where(W > F32(1e-10), q, oldLab).astype(F32)
        nl = np.minimum(np.maximum(nl, F32(0)), F32(1)).astype(F32)
        new[lst] = nl
        changedMask = np.abs((nl - oldLab).astype(F32)) > d
        # markNeighborsKernel (:187-217) into a cleared next array
        nxt = np.zeros(N, np.int32)
        for u in lst[changedMask]:
            for e in range(rowPtr[u], rowPtr[u + 1]):
                v = colInd[e]
                if v >= 0 and status[v] != 1 and gtClass[v] < 0:
                    nxt[v] = 1
            if gtClass[u] < 0:
                nxt[u] = 1
        old, new = new, old              # :308-310
        aff = nxt                        # :314-316
        iterationsUsed = it + 1
        if not changedMask.any():
            break
    if old is not bufA:                  # :330-334
        bufA[:] = old
    bufA[gtClass == 0] = F32(0)          # enforceGroundTruthKernel
    bufA[gtClass == 1] = F32(1)
    return bufA, bufB, aff, iterationsUsed, np.array(sizes, np.int32)



def block_row_sums(rowPtr, colInd, values, status, old, lst):
    """W and S of rows `lst` as a 256-thread CUDA block sums them
    (iterativePropagation.cu:122-150, validation.cu:78-105): thread t adds
    row positions t, t+256, ... in order into its slot, skipping negative
    columns and DELETED neighbours, then s[t] += s[t+stride] for stride =
    128 .. 1, all in float32."""
    starts = rowPtr[lst]
    lens = rowPtr[lst + 1] - starts
    rounds = int((lens.max() + 255) // 256) if lens.size else 0
    sW = np.zeros((lst.size, 256), F32)
    sS = np.zeros((lst.size, 256), F32)
    for r in range(rounds):          # sequential per thread: e = start + t + 256 r
        pos = r * 256 + np.arange(256)[None, :]
        valid = pos < lens[:, None]
        e = np.where(valid, starts[:, None] + pos, 0)
        v = colInd[e]
        ok = valid & (v >= 0)
        vv = np.where(ok, v, 0)
        ok &= status[vv] != 1
        w = np.where(ok, values[e], F32(0))
        fv = np.where(ok, old[vv], F32(0))
        prod = (w * fv).astype(F32)
        # a skipped element leaves the slot unchanged
        sW = np.where(ok, (sW + w).astype(F32), sW)
        sS = np.where(ok, (sS + prod).astype(F32), sS)
    stride = 128
    while stride > 0:                # shared-memory tree
        sW[:, :stride] = (sW[:, :stride] + sW[:, stride:2 * stride]).astype(F32)
        sS[:, :stride] = (sS[:, :stride] + sS[:, stride:2 * stride]).astype(F32)
        stride //= 2
    return sW[:, 0], sS[:, 0]


def ref_irlp(rowPtr, colInd, values, labels, status, gtClass, delta, maxIterations):
    """launchIrLP (validation.cu:144-205) with irLPKernel (:47-127) and
    irLPEnforceGTKernel (:132-139): returns (labels, iterationsUsed)."""
    rowPtr = np.asarray(rowPtr, np.int64)
    colInd = np.asarray(colInd, np.int64)
    values = np.asarray(values, F32)
    status = np.asarray(status)
    gtClass = np.asarray(gtClass)
    N = rowPtr.shape[0] - 1
    bufA = np.array(labels, F32)            # d_labels
    bufB = bufA.copy()                      # d_labelsBuffer (:157-159)
    old, new = bufA, bufB
    d = F32(delta)
    iterationsUsed = 0
    for it in range(maxIterations):
        # irLPKernel: GT and non-active blocks copy the old label (:56-70)
        new[:] = old
        lst = np.nonzero((gtClass < 0) & (status == 0))[0]
        W, S = block_row_sums(rowPtr, colInd, values, status, old, lst)
        oldLab = old[lst]
        with np.errstate(divide="ignore", invalid="ignore"):
            q = (S / W).astype(F32)
        nl = np.where(W > F32(1e-10), q, oldLab).astype(F32)
        nl = np.minimum(np.maximum(nl, F32(0)), F32(1)).astype(F32)
        new[lst] = nl
        changed = bool((np.abs((nl - oldLab).astype(F32)) > d).any())
        new[gtClass == 0] = F32(0)          # irLPEnforceGTKernel
        new[gtClass == 1] = F32(1)
        old, new = new, old                 # :185-189
        iterationsUsed = it + 1
        if not changed:                     # :192-196
            break
    if old is not bufA:                     # :199-203
        bufA[:] = old
    return bufA, iterationsUsed

def ref_mark_deletions(rowPtr, colInd, status, isAffected, deleted, numNodes):
    colInd = np.array(colInd, np.int32)
    status = np.array(status, np.int32)
    aff = np.array(isAffected, np.int32)
    for u in deleted:                    # markDeletionsKernel (:58-83)
        if u < 0 or u >= numNodes:
            continue
        for e in range(rowPtr[u], rowPtr[u + 1]):
            v = colInd[e]
            if 0 <= v < numNodes:
                aff[v] = max(aff[v], 1)
            colInd[e] = -1
        status[u] = 1
    for u in range(numNodes):            # cleanIncomingEdgesKernel (:92-111)
        if status[u] == 1:
            continue
        for e in range(rowPtr[u], rowPtr[u + 1]):
            v = colInd[e]
            if 0 <= v < numNodes and status[v] == 1:
                colInd[e] = -1
    return colInd, status, aff


def ref_active_csr(fullRowPtr, fullColInd, fullValues, hostStatus, batchMap, b):
    """updateHostCSR's rebuild (dynLP.cu:134-166)."""
    N = len(fullRowPtr) - 1
    rowPtr = [0]
    cols, vals = [], []

    def active(x):
        return hostStatus[x] != 1 and 0 <= batchMap[x] <= b

    for u in range(N):
        if active(u):
            for e in range(fullRowPtr[u], fullRowPtr[u + 1]):
                v = fullColInd[e]
                if v >= 0 and active(v):
                    cols.append(v)
                    vals.append(fullValues[e])
        rowPtr.append(len(cols))
    return (np.array(rowPtr, np.int32), np.array(cols, np.int32),
            np.array(vals, np.float32))
