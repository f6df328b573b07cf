"""ctypes wrapper of the CPU oracle (oracle/libdynlp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline --
never by the product (labelpropcuda_amd/).

The oracle restates the reference's runDynLP (src/dynLP.cu:175-587) and its
step launchers literally in C; see dynlp_oracle.h for the three forced
deviations (CC over G', fixed label-init summation order, IEEE division).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdynlp_oracle.so")

i32p = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)

_lib = None


def build() -> None:
    """Compile the oracle library (make -C oracle)."""
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    for pfx, rp in (("orc32_", f32p), ("orc64_", f64p)):
        sig = {
            "mark_deletions": (None, [i32p, i32p, i32p, i32p, i32p, C.c_int, C.c_int]),
            "sparsify": (C.c_int, [i32p, i32p, f32p, C.c_int, C.c_int, i32p, C.c_float, i32p,
                                   i32p, i32p, f32p]),
            "connected_components": (C.c_int, [i32p, i32p, C.c_int, i32p, C.c_int, i32p,
                                               C.c_int, i32p]),
            "label_initialization": (C.c_int, [i32p, i32p, f32p, rp, i32p, C.c_int, i32p,
                                               C.c_int, i32p]),
            "propagate": (C.c_int, [i32p, i32p, f32p, rp, rp, i32p, i32p, i32p, C.c_int,
                                    C.c_float, C.c_int, i32p, C.c_int, i32p]),
            "irlp": (C.c_int, [i32p, i32p, f32p, rp, i32p, i32p, C.c_int, C.c_float, C.c_int]),
            "set_threads": (None, [C.c_int, C.c_int]),
            "create": (C.c_void_p, [C.c_int, i32p, i32p, f32p, f32p, C.c_int, i64p, i32p, i64p,
                                    i32p, f32p, i64p, i32p, i64p, i32p, i32p, f32p, C.c_float,
                                    C.c_float, C.c_int]),
            "destroy": (None, [C.c_void_p]),
            "tau": (C.c_float, [C.c_void_p]),
            "run_batch": (C.c_int, [C.c_void_p, C.c_int]),
            "nnz": (C.c_int, [C.c_void_p]),
            "get_csr": (None, [C.c_void_p, i32p, i32p, f32p]),
            "get_affected": (None, [C.c_void_p, i32p]),
            "get_sub": (C.c_int, [C.c_void_p, i32p, i32p, f32p]),
            "get_component_ids": (C.c_int, [C.c_void_p, i32p]),
            "get_init_labels": (None, [C.c_void_p, rp]),
            "get_labels": (None, [C.c_void_p, rp]),
            "get_frontier_sizes": (C.c_int, [C.c_void_p, i32p, C.c_int]),
            "get_status": (None, [C.c_void_p, i32p, i32p, i32p]),
            "get_stats": (None, [C.c_void_p, i32p, i64p, f32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, pfx + name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


def set_threads(threads: int, minFrontier: int = 1024) -> None:
    """Worker threads of the propagation loop (1 = the literal sequential
    restatement; more = the frontier-parallel variant, same results) for
    frontiers of at least minFrontier vertices."""
    lib = load()
    lib.orc32_set_threads(int(threads), int(minFrontier))
    lib.orc64_set_threads(int(threads), int(minFrontier))


def _p(a: np.ndarray):
    t = {np.dtype(np.int32): C.c_int, np.dtype(np.int64): C.c_int64,
         np.dtype(np.float32): C.c_float, np.dtype(np.float64): C.c_double}[a.dtype]
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """CPU runDynLP over an ``Inputs`` (labelpropcuda_amd.dynlp.Inputs or any
    object with the same arrays); FP32 by default, FP64 twin with fp64=True."""

    def __init__(self, inputs, tau: float = -1.0, delta: float = 1e-4,
                 maxIterations: int = 100000, fp64: bool = False):
        self.lib = load()
        self.pfx = "orc64_" if fp64 else "orc32_"
        self.real = np.float64 if fp64 else np.float32
        self.inputs = inputs
        self.N = int(inputs.rowPtr.shape[0] - 1)
        self.B = int(inputs.insOff.shape[0] - 1)
        # keep every array alive and contiguous for the C side
        self._arrays = [np.ascontiguousarray(a) for a in (
            inputs.rowPtr, inputs.colInd, inputs.values, inputs.trueLabels, inputs.insOff,
            inputs.ins, inputs.gtOff, inputs.gtIds, inputs.gtLab, inputs.delOff, inputs.dels,
            inputs.eOff, inputs.eSrc, inputs.eDst, inputs.eW)]
        a = self._arrays
        self.h = self._fn("create")(self.N, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), self.B,
                                    _p(a[4]), _p(a[5]), _p(a[6]), _p(a[7]), _p(a[8]), _p(a[9]),
                                    _p(a[10]), _p(a[11]), _p(a[12]), _p(a[13]), _p(a[14]),
                                    tau, delta, maxIterations)
        self.maxIterations = maxIterations

    def _fn(self, name):
        return getattr(self.lib, self.pfx + name)

    def close(self):
        if self.h:
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tau(self) -> float:
        return float(self._fn("tau")(self.h))

    def run_batch(self, b: int) -> None:
        rc = self._fn("run_batch")(self.h, b)
        if rc != 0:
            raise RuntimeError(f"oracle run_batch({b}) failed with {rc}")

    def run_all(self) -> None:
        for b in range(self.B):
            self.run_batch(b)

    def csr(self):
        nnz = int(self._fn("nnz")(self.h))
        rp = np.zeros(self.N + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        vv = np.zeros(max(nnz, 1), np.float32)
        self._fn("get_csr")(self.h, _p(rp), _p(ci), _p(vv))
        return rp, ci[:nnz], vv[:nnz]

    def affected(self) -> np.ndarray:
        out = np.zeros(self.N, np.int32)
        self._fn("get_affected")(self.h, _p(out))
        return out

    def subgraph(self):
        n = int(self._fn("get_sub")(self.h, None, None, None))
        s = np.zeros(max(n, 1), np.int32)
        d = np.zeros(max(n, 1), np.int32)
        w = np.zeros(max(n, 1), np.float32)
        self._fn("get_sub")(self.h, _p(s), _p(d), _p(w))
        return s[:n], d[:n], w[:n]

    def component_ids(self):
        out = np.zeros(self.N, np.int32)
        n = int(self._fn("get_component_ids")(self.h, _p(out)))
        return out, n

    def init_labels(self) -> np.ndarray:
        out = np.zeros(self.N, self.real)
        self._fn("get_init_labels")(self.h, _p(out))
        return out

    def labels(self) -> np.ndarray:
        out = np.zeros(self.N, self.real)
        self._fn("get_labels")(self.h, _p(out))
        return out

    def frontier_sizes(self) -> np.ndarray:
        cap = min(self.maxIterations, 1 << 20) + 1
        out = np.zeros(cap, np.int32)
        n = int(self._fn("get_frontier_sizes")(self.h, _p(out), cap))
        return out[:min(n, cap)]

    def status(self):
        """(device status during the last batch, host status, gtClass)."""
        d = np.zeros(self.N, np.int32)
        h = np.zeros(self.N, np.int32)
        g = np.zeros(self.N, np.int32)
        self._fn("get_status")(self.h, _p(d), _p(h), _p(g))
        return d, h, g

    def stats(self) -> dict:
        s5 = np.zeros(5, np.int32)
        upd = np.zeros(1, np.int64)
        acc = np.zeros(1, np.float32)
        self._fn("get_stats")(self.h, _p(s5), _p(upd), _p(acc))
        return {"iterations": int(s5[0]), "numSubEdges": int(s5[1]),
                "numComponents": int(s5[2]), "numMarked": int(s5[3]),
                "frontier0": int(s5[4]), "updates": int(upd[0]), "accuracy": float(acc[0])}


# ---- step restatements on numpy arrays (each returns new arrays) ---------

def mark_deletions(rowPtr, colInd, status, isAffected, deleted, numNodes, fp64=False):
    lib = load()
    pfx = "orc64_" if fp64 else "orc32_"
    colInd = np.array(colInd, np.int32)
    status = np.array(status, np.int32)
    isAffected = np.array(isAffected, np.int32)
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    deleted = np.ascontiguousarray(deleted, np.int32)
    getattr(lib, pfx + "mark_deletions")(_p(rowPtr), _p(colInd), _p(status), _p(isAffected),
                                         _p(deleted), int(deleted.shape[0]), numNodes)
    return colInd, status, isAffected


def sparsify(src, dst, w, numInserted, isNew, tau, isAffected):
    lib = load()
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    w = np.ascontiguousarray(w, np.float32)
    isNew = np.ascontiguousarray(isNew, np.int32)
    isAffected = np.array(isAffected, np.int32)
    n = int(src.shape[0])
    ss = np.zeros(max(2 * n, 1), np.int32)
    sd = np.zeros(max(2 * n, 1), np.int32)
    sw = np.zeros(max(2 * n, 1), np.float32)
    cnt = lib.orc32_sparsify(_p(src), _p(dst), _p(w), n, numInserted, _p(isNew), tau,
                             _p(isAffected), _p(ss), _p(sd), _p(sw))
    return ss[:cnt], sd[:cnt], sw[:cnt], isAffected


def connected_components(src, dst, inserted, numNodes):
    lib = load()
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    inserted = np.ascontiguousarray(inserted, np.int32)
    comp = np.full(numNodes, -7, np.int32)
    nc = np.zeros(1, np.int32)
    rc = lib.orc32_connected_components(_p(src), _p(dst), int(src.shape[0]), _p(inserted),
                                        int(inserted.shape[0]), _p(comp), numNodes, _p(nc))
    if rc != 0:
        raise ValueError("edge endpoint is not an inserted vertex")
    return comp, int(nc[0])


def label_initialization(rowPtr, colInd, values, labels, componentId, numComponents, inserted,
                         gtClass, fp64=False):
    lib = load()
    pfx = "orc64_" if fp64 else "orc32_"
    real = np.float64 if fp64 else np.float32
    labels = np.array(labels, real)
    args = [np.ascontiguousarray(a, t) for a, t in ((rowPtr, np.int32), (colInd, np.int32),
                                                    (values, np.float32),
                                                    (componentId, np.int32),
                                                    (inserted, np.int32), (gtClass, np.int32))]
    rc = getattr(lib, pfx + "label_initialization")(
        _p(args[0]), _p(args[1]), _p(args[2]), _p(labels), _p(args[3]), numComponents,
        _p(args[4]), int(args[4].shape[0]), _p(args[5]))
    if rc != 0:
        raise ValueError("componentId out of range")
    return labels


def propagate(rowPtr, colInd, values, labels, status, gtClass, isAffected, delta,
              maxIterations, fp64=False):
    """Returns (labels, labelsBuffer, isAffected, iterationsUsed, frontierSizes)."""
    lib = load()
    pfx = "orc64_" if fp64 else "orc32_"
    real = np.float64 if fp64 else np.float32
    labels = np.array(labels, real)
    buf = np.zeros_like(labels)
    isAffected = np.array(isAffected, np.int32)
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    colInd = np.ascontiguousarray(colInd, np.int32)
    values = np.ascontiguousarray(values, np.float32)
    status = np.ascontiguousarray(status, np.int32)
    gtClass = np.ascontiguousarray(gtClass, np.int32)
    N = int(rowPtr.shape[0] - 1)
    cap = min(maxIterations, 1 << 20) + 1
    fs = np.zeros(cap, np.int32)
    nfs = np.zeros(1, np.int32)
    it = getattr(lib, pfx + "propagate")(_p(rowPtr), _p(colInd), _p(values), _p(labels), _p(buf),
                                         _p(status), _p(gtClass), _p(isAffected), N, delta,
                                         maxIterations, _p(fs), cap, _p(nfs))
    return labels, buf, isAffected, int(it), fs[:min(int(nfs[0]), cap)]


def irlp(rowPtr, colInd, values, labels, status, gtClass, delta, maxIterations, fp64=False):
    """IrLP full recompute (validation.cu:47-213): returns (labels, iterationsUsed)."""
    lib = load()
    pfx = "orc64_" if fp64 else "orc32_"
    real = np.float64 if fp64 else np.float32
    labels = np.array(labels, real)
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    it = getattr(lib, pfx + "irlp")(_p(rowPtr), _p(np.ascontiguousarray(colInd, np.int32)),
                                     _p(np.ascontiguousarray(values, np.float32)), _p(labels),
                                     _p(np.ascontiguousarray(status, np.int32)),
                                     _p(np.ascontiguousarray(gtClass, np.int32)),
                                     int(rowPtr.shape[0] - 1), delta, maxIterations)
    return labels, int(it)


def irlp_final_state(inputs, numBatches=None):
    """The graph IrLP runs on (main.cu:124-187): status after replaying every
    batch (inserted -> ACTIVE, then deleted -> DELETED, then GT), labels 0.5
    except GT, and the CSR of non-deleted inserted vertices filtered to
    non-deleted inserted neighbours.  Returns (rowPtr, colInd, values,
    status, gtClass, labels)."""
    N = inputs.numNodes
    B = inputs.numBatches if numBatches is None else numBatches
    status = np.full(N, 1, np.int32)   # DELETED
    gt = np.full(N, -1, np.int32)
    labels = np.full(N, 0.5, np.float32)
    batchOf = np.full(N, -1, np.int32)
    for b in range(B):
        bt = inputs.batch(b)
        status[bt["ins"]] = 0
        status[bt["dels"]] = 1
        status[bt["gtIds"]] = 2
        gt[bt["gtIds"]] = (bt["gtLab"] >= 0.5).astype(np.int32)
        labels[bt["gtIds"]] = bt["gtLab"]
        batchOf[bt["ins"]] = b
    keep = (status != 1) & (batchOf >= 0)
    rp, ci, vv = inputs.rowPtr, inputs.colInd, inputs.values
    rowPtr = [0]
    cols, vals = [], []
    for u in range(N):
        if keep[u]:
            seg = slice(rp[u], rp[u + 1])
            c = ci[seg]
            ok = (c >= 0) & keep[np.maximum(c, 0)]
            cols.append(c[ok])
            vals.append(vv[seg][ok])
            rowPtr.append(rowPtr[-1] + int(ok.sum()))
        else:
            rowPtr.append(rowPtr[-1])
    colInd = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
    values = np.concatenate(vals).astype(np.float32) if vals else np.zeros(0, np.float32)
    return np.array(rowPtr, np.int32), colInd, values, status, gt, labels
