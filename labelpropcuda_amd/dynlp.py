"""Python host layer over libdynlp's C-ABI (plumbing for tests and bench).

Mirrors the reference's driver objects: ``Inputs`` holds what
``generateSparseGraph`` + ``generateBatches`` produce (CSRGraph, trueLabels,
std::vector<BatchData> flattened), ``run_dynlp`` is ``runDynLP``
(dynLP.h:69-74), and ``Context`` is the device-resident form of the same batch
loop with per-batch stats and an optional intermediate-state trace.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, load


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Inputs:
    """Full graph, true labels and the flattened batch stream."""

    rowPtr: np.ndarray
    colInd: np.ndarray
    values: np.ndarray
    trueLabels: np.ndarray
    insOff: np.ndarray
    ins: np.ndarray
    gtOff: np.ndarray
    gtIds: np.ndarray
    gtLab: np.ndarray
    delOff: np.ndarray
    dels: np.ndarray
    eOff: np.ndarray
    eSrc: np.ndarray
    eDst: np.ndarray
    eW: np.ndarray
    _keep: list = field(default_factory=list, repr=False)

    @property
    def numNodes(self) -> int:
        return int(self.rowPtr.shape[0] - 1)

    @property
    def numEdges(self) -> int:
        return int(self.colInd.shape[0])

    @property
    def numBatches(self) -> int:
        return int(self.insOff.shape[0] - 1)

    def batch(self, b: int) -> dict:
        s = slice
        return {
            "ins": self.ins[s(self.insOff[b], self.insOff[b + 1])],
            "gtIds": self.gtIds[s(self.gtOff[b], self.gtOff[b + 1])],
            "gtLab": self.gtLab[s(self.gtOff[b], self.gtOff[b + 1])],
            "dels": self.dels[s(self.delOff[b], self.delOff[b + 1])],
            "eSrc": self.eSrc[s(self.eOff[b], self.eOff[b + 1])],
            "eDst": self.eDst[s(self.eOff[b], self.eOff[b + 1])],
            "eW": self.eW[s(self.eOff[b], self.eOff[b + 1])],
        }

    _ARRAYS = ("rowPtr", "colInd", "values", "trueLabels", "insOff", "ins", "gtOff", "gtIds",
               "gtLab", "delOff", "dels", "eOff", "eSrc", "eDst", "eW")

    def save(self, path: str) -> None:
        """Binary snapshot of the generator's output (uncompressed .npz)."""
        np.savez(path, **{k: getattr(self, k) for k in self._ARRAYS})

    @classmethod
    def load(cls, path: str) -> "Inputs":
        with np.load(path, allow_pickle=False) as z:
            return cls(**{k: np.ascontiguousarray(z[k]) for k in cls._ARRAYS})

    def graph_view(self) -> _lib.GraphView:
        g = _lib.GraphView()
        g.numNodes = self.numNodes
        g.numEdges = self.numEdges
        g.rowPtr = _ptr(self.rowPtr, C.c_int)
        g.colInd = _ptr(self.colInd, C.c_int)
        g.values = _ptr(self.values, C.c_float)
        return g

    def batch_view(self) -> _lib.BatchView:
        v = _lib.BatchView()
        v.numBatches = self.numBatches
        v.insertedOffsets = _ptr(self.insOff, C.c_int64)
        v.insertedNodeIds = _ptr(self.ins, C.c_int)
        v.groundTruthOffsets = _ptr(self.gtOff, C.c_int64)
        v.groundTruthNodeIds = _ptr(self.gtIds, C.c_int)
        v.groundTruthLabels = _ptr(self.gtLab, C.c_float)
        v.deletedOffsets = _ptr(self.delOff, C.c_int64)
        v.deletedNodeIds = _ptr(self.dels, C.c_int)
        v.edgeOffsets = _ptr(self.eOff, C.c_int64)
        v.newEdgeSrc = _ptr(self.eSrc, C.c_int)
        v.newEdgeDst = _ptr(self.eDst, C.c_int)
        v.newEdgeWeight = _ptr(self.eW, C.c_float)
        return v


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def generate(totalNodes: int, avgDegree: int, numBatches: int, seed: int = 42,
             gtFraction: float = 0.01, deletedFraction: float = 0.09) -> Inputs:
    """generateSparseGraph + generateBatches (same RNG sequence as the reference)."""
    lib = load()
    h = C.c_void_p()
    check(lib.dlp_inputs_generate(C.byref(h), totalNodes, avgDegree, numBatches, seed,
                                  gtFraction, deletedFraction))
    return _inputs_from_handle(lib, h)


def generate_stream(totalNodes: int, avgDegree: int, batchSizes, seed: int = 42,
                    gtFraction: float = 0.01, deletedFraction: float = 0.09) -> Inputs:
    """The reference graph generator with batches of the given sizes
    (batchSizes[0] is ignored: batch 0 is the GT seed batch)."""
    lib = load()
    sizes = np.ascontiguousarray(batchSizes, np.int32)
    h = C.c_void_p()
    check(lib.dlp_inputs_generate_stream(C.byref(h), totalNodes, avgDegree,
                                         _ptr(sizes, C.c_int), len(sizes), seed, gtFraction,
                                         deletedFraction))
    return _inputs_from_handle(lib, h)


def write_graph(inputs: Inputs, prefix: str) -> None:
    """The generator's CSR text dump of `inputs`' full graph and true labels
    (<prefix>RowPtr.txt, ColInd.txt, Values.txt, TrueLabels.txt)."""
    g = inputs.graph_view()
    check(load().dlp_write_graph_files(C.byref(g), _ptr(inputs.trueLabels, C.c_float),
                                       prefix.encode()))


def from_graph_files(prefix: str, numBatches: int, seed: int = 42, gtFraction: float = 0.01,
                     deletedFraction: float = 0.09) -> Inputs:
    """A graph read back from the CSR text dump, with the reference batch
    generator run over it."""
    lib = load()
    h = C.c_void_p()
    check(lib.dlp_inputs_from_graph_files(C.byref(h), prefix.encode(), numBatches, seed,
                                          gtFraction, deletedFraction))
    return _inputs_from_handle(lib, h)


def _inputs_from_handle(lib, h) -> Inputs:
    try:
        g = _lib.GraphView()
        tl = _lib.f32p()
        check(lib.dlp_inputs_graph(h, C.byref(g), C.byref(tl)))
        bv = _lib.BatchView()
        check(lib.dlp_inputs_batches(h, C.byref(bv)))
        N, E, B = g.numNodes, g.numEdges, bv.numBatches
        insOff = _copy(bv.insertedOffsets, B + 1, np.int64)
        gtOff = _copy(bv.groundTruthOffsets, B + 1, np.int64)
        delOff = _copy(bv.deletedOffsets, B + 1, np.int64)
        eOff = _copy(bv.edgeOffsets, B + 1, np.int64)
        return Inputs(
            rowPtr=_copy(g.rowPtr, N + 1, np.int32),
            colInd=_copy(g.colInd, E, np.int32),
            values=_copy(g.values, E, np.float32),
            trueLabels=_copy(tl, N, np.float32),
            insOff=insOff, ins=_copy(bv.insertedNodeIds, int(insOff[-1]), np.int32),
            gtOff=gtOff, gtIds=_copy(bv.groundTruthNodeIds, int(gtOff[-1]), np.int32),
            gtLab=_copy(bv.groundTruthLabels, int(gtOff[-1]), np.float32),
            delOff=delOff, dels=_copy(bv.deletedNodeIds, int(delOff[-1]), np.int32),
            eOff=eOff, eSrc=_copy(bv.newEdgeSrc, int(eOff[-1]), np.int32),
            eDst=_copy(bv.newEdgeDst, int(eOff[-1]), np.int32),
            eW=_copy(bv.newEdgeWeight, int(eOff[-1]), np.float32),
        )
    finally:
        lib.dlp_inputs_free(h)


def make_config(inputs: Inputs, tau: float = -1.0, delta: float = 1e-4,
                maxIterations: int = 100000, verbose: bool = False,
                numBatches: int | None = None) -> _lib.Config:
    cfg = _lib.Config()
    load().dlp_config_default(C.byref(cfg))
    cfg.totalNodes = inputs.numNodes
    cfg.numBatches = inputs.numBatches if numBatches is None else numBatches
    cfg.tau = tau
    cfg.delta = delta
    cfg.maxIterations = maxIterations
    cfg.verbose = 1 if verbose else 0
    cfg.validate = 0
    return cfg


def run_dynlp(inputs: Inputs, cfg: _lib.Config):
    """runDynLP: returns (finalLabels, batchTimesMs, batchAccuracy)."""
    lib = load()
    N, B = inputs.numNodes, cfg.numBatches
    labels = np.zeros(N, np.float32)
    times = np.zeros(B, np.float32)
    acc = np.zeros(B, np.float32)
    g, bv = inputs.graph_view(), inputs.batch_view()
    check(lib.dlp_run_dynlp(C.byref(cfg), C.byref(g), _ptr(inputs.trueLabels, C.c_float),
                            C.byref(bv), _ptr(labels, C.c_float), _ptr(times, C.c_float),
                            _ptr(acc, C.c_float)))
    return labels, times, acc


def irlp_tuning(**kw) -> _lib.IrlpTuning:
    """IrLP launch shape: library defaults, then the given fields."""
    t = _lib.IrlpTuning()
    load().dlp_irlp_tuning_default(C.byref(t))
    for k, v in kw.items():
        setattr(t, k, int(v))
    return t


def irlp(rowPtr, colInd, values, labels, status, gtClass, delta: float = 1e-4,
         maxIterations: int = 100000, tuning: _lib.IrlpTuning | None = None):
    """launchIrLP (validation.cu:144-205) on host arrays, through dlp_irlp_ex:
    returns (labels, iterationsUsed, loopMs)."""
    lib = load()
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    N = int(rowPtr.shape[0] - 1)
    dev = [_lib.DeviceArray(np.ascontiguousarray(a, t)) for a, t in
           ((rowPtr, np.int32), (colInd, np.int32), (values, np.float32),
            (labels, np.float32), (status, np.int32), (gtClass, np.int32))]
    it = C.c_int(0)
    ms = C.c_float(0.0)
    try:
        check(lib.dlp_irlp_ex(dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, dev[4].ptr,
                              dev[5].ptr, N, C.c_float(delta), int(maxIterations), C.byref(it),
                              None, C.byref(tuning) if tuning is not None else None,
                              C.byref(ms)))
        out = dev[3].numpy()
    finally:
        for d in dev:
            d.free()
    return out, it.value, ms.value


def irlp_validation(rowPtr, colInd, values, status, gtClass, labels, delta: float = 1e-4,
                    maxIterations: int = 100000):
    """runIrLPValidation (validation.cu:210-276) over host arrays: returns
    (labels, iterationsUsed, skipped)."""
    lib = load()
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    colInd = np.ascontiguousarray(colInd, np.int32)
    values = np.ascontiguousarray(values, np.float32)
    status = np.ascontiguousarray(status, np.int32)
    gtClass = np.ascontiguousarray(gtClass, np.int32)
    out = np.array(labels, np.float32)
    g = _lib.GraphView()
    g.numNodes = int(rowPtr.shape[0] - 1)
    g.numEdges = int(colInd.shape[0])
    g.rowPtr = _ptr(rowPtr, C.c_int)
    g.colInd = _ptr(colInd, C.c_int)
    g.values = _ptr(values, C.c_float)
    it, skipped = C.c_int(0), C.c_int(0)
    check(lib.dlp_irlp_validation(C.byref(g), _ptr(status, C.c_int), _ptr(gtClass, C.c_int),
                                  _ptr(out, C.c_float), g.numNodes, C.c_float(delta),
                                  int(maxIterations), C.byref(it), C.byref(skipped)))
    return out, it.value, bool(skipped.value)


def irlp_baseline(inputs: Inputs, cfg: _lib.Config, tuning: _lib.IrlpTuning | None = None):
    """The --validate step (main.cu:124-217): IrLP on the final state of the
    batch stream.  Returns (irlpLabels, {iterations, skipped, activeEdges,
    timeMs, loopMs, accuracy})."""
    lib = load()
    labels = np.zeros(inputs.numNodes, np.float32)
    res = _lib.IrlpResult()
    g, bv = inputs.graph_view(), inputs.batch_view()
    check(lib.dlp_irlp_baseline(C.byref(cfg), C.byref(g), _ptr(inputs.trueLabels, C.c_float),
                                C.byref(bv), C.byref(tuning) if tuning is not None else None,
                                _ptr(labels, C.c_float), C.byref(res)))
    return labels, {f: getattr(res, f) for f, _ in res._fields_}


class TraceBuffers:
    """Host buffers for dlp_trace (the intermediate state of one batch)."""

    def __init__(self, inputs: Inputs, maxFrontierSizes: int = 1 << 16):
        N, E = inputs.numNodes, inputs.numEdges
        maxE = int(np.max(np.diff(inputs.eOff))) if inputs.numBatches else 0
        self.rowPtr = np.zeros(N + 1, np.int32)
        self.colInd = np.zeros(max(E, 1), np.int32)
        self.values = np.zeros(max(E, 1), np.float32)
        self.affected = np.zeros(N, np.int32)
        self.subSrc = np.zeros(max(2 * maxE, 1), np.int32)
        self.subDst = np.zeros(max(2 * maxE, 1), np.int32)
        self.subW = np.zeros(max(2 * maxE, 1), np.float32)
        self.compId = np.zeros(N, np.int32)
        self.initLabels = np.zeros(N, np.float32)
        self.frontier = np.zeros(maxFrontierSizes, np.int32)
        t = _lib.Trace()
        t.csrRowPtr = _ptr(self.rowPtr, C.c_int)
        t.csrColInd = _ptr(self.colInd, C.c_int)
        t.csrValues = _ptr(self.values, C.c_float)
        t.affected_N = _ptr(self.affected, C.c_int)
        t.subSrc = _ptr(self.subSrc, C.c_int)
        t.subDst = _ptr(self.subDst, C.c_int)
        t.subW = _ptr(self.subW, C.c_float)
        t.componentId_N = _ptr(self.compId, C.c_int)
        t.initLabels_N = _ptr(self.initLabels, C.c_float)
        t.frontierSizes = _ptr(self.frontier, C.c_int)
        t.maxFrontierSizes = maxFrontierSizes
        self.t = t

    def snapshot(self) -> dict:
        nnz = int(self.rowPtr[-1])
        ns = self.t.numSubEdges
        nf = min(self.t.numFrontierSizes, self.frontier.shape[0])
        return {
            "rowPtr": self.rowPtr.copy(),
            "colInd": self.colInd[:nnz].copy(),
            "values": self.values[:nnz].copy(),
            "affected": self.affected.copy(),
            "subSrc": self.subSrc[:ns].copy(),
            "subDst": self.subDst[:ns].copy(),
            "subW": self.subW[:ns].copy(),
            "numSubEdges": ns,
            "numComponents": self.t.numComponents,
            "compId": self.compId.copy(),
            "initLabels": self.initLabels.copy(),
            "frontierSizes": self.frontier[:nf].copy(),
            "iterations": self.t.numFrontierSizes,
        }


class Context:
    """Device-resident graph + batch stream (dlp_context_*)."""

    def __init__(self, inputs: Inputs, cfg: _lib.Config):
        self._lib = load()
        self.inputs = inputs
        self.cfg = cfg
        self._h = C.c_void_p()
        g, bv = inputs.graph_view(), inputs.batch_view()
        check(self._lib.dlp_context_create(C.byref(self._h), C.byref(cfg), C.byref(g),
                                           _ptr(inputs.trueLabels, C.c_float), C.byref(bv),
                                           None))
        self._trace = None

    def close(self):
        if self._h:
            self._lib.dlp_context_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tau(self) -> float:
        t = C.c_float()
        check(self._lib.dlp_context_tau(self._h, C.byref(t)))
        return t.value

    def device_bytes(self) -> int:
        b = C.c_int64()
        check(self._lib.dlp_context_device_bytes(self._h, C.byref(b)))
        return b.value

    def set_tuning(self, groupLanes=0, blocks=0, chunk=0, localThreshold=0, localLdsBytes=0,
                   localShape=0, localOwnerAfter=0, waveRowMin=0, blockRowMin=0,
                   maskOwnerMax=0):
        """0 keeps a setting; localThreshold < 0 disables the LDS-resident
        propagation kernel; localShape is 1-based; localOwnerAfter -1 runs
        the LDS-resident kernel with queue iterations only; rows longer than
        waveRowMin / blockRowMin go to a wave / a work-group (blockRowMin -1:
        never); owner mode on frontiers of at most maskOwnerMax vertices runs in
        the register-mask kernels (-1: never)."""
        t = _lib.Tuning(groupLanes, blocks, chunk, localThreshold, localLdsBytes, localShape,
                        localOwnerAfter, waveRowMin, blockRowMin, maskOwnerMax)
        check(self._lib.dlp_context_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = _lib.Tuning()
        check(self._lib.dlp_context_get_tuning(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_}

    def enable_trace(self, maxFrontierSizes: int = 1 << 16) -> TraceBuffers:
        self._trace = TraceBuffers(self.inputs, maxFrontierSizes)
        check(self._lib.dlp_context_set_trace(self._h, C.byref(self._trace.t)))
        return self._trace

    def disable_trace(self):
        check(self._lib.dlp_context_set_trace(self._h, None))
        self._trace = None

    def set_profiling(self, level=True):
        """0/False off, 1 event timing per local launch and per multi-block
        chunk, 2/True per-launch events plus work counters."""
        lv = 2 if level is True else (0 if level is False else int(level))
        check(self._lib.dlp_context_set_profiling(self._h, lv))

    def prop_profile(self) -> dict:
        p = _lib.PropProfile()
        check(self._lib.dlp_context_prop_profile(self._h, C.byref(p)))
        return {name: getattr(p, name) for name, _ in p._fields_}

    def row_update(self, mode: int = -1) -> bool:
        """Sets the CSR update's row choice (0 by batch size, 1 every row, 2
        only the rows the batch can change on a symmetric graph; -1 keeps
        it); returns whether the full graph is symmetric."""
        sym = C.c_int()
        check(self._lib.dlp_context_row_update(self._h, mode, C.byref(sym)))
        return bool(sym.value)

    def reset(self):
        check(self._lib.dlp_context_reset(self._h))

    def run_batch(self, b: int):
        check(self._lib.dlp_context_run_batch(self._h, b))

    def run_all(self):
        check(self._lib.dlp_context_run_all(self._h))

    def synchronize(self):
        check(self._lib.dlp_context_synchronize(self._h))

    def labels(self) -> np.ndarray:
        out = np.zeros(self.inputs.numNodes, np.float32)
        check(self._lib.dlp_context_read_labels(self._h, _ptr(out, C.c_float)))
        return out

    def stats(self, b: int) -> dict:
        s = _lib.BatchStats()
        check(self._lib.dlp_context_batch_stats(self._h, b, C.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def autotune(self, trialBatches: int = 0):
        """Times candidate launch shapes on the first trialBatches batches
        (all when 0), keeps the fastest and stores it in the process's tuning
        table; returns (tuning, ms of the best trial)."""
        t = _lib.Tuning()
        ms = C.c_float()
        check(self._lib.dlp_context_autotune(self._h, trialBatches, C.byref(t), C.byref(ms)))
        return {name: getattr(t, name) for name, _ in t._fields_}, ms.value

    def shape_key(self):
        """(shape key, tuning source: 'heuristic' | 'table' | 'autotune')."""
        k = _lib.ShapeKey()
        src = C.c_int()
        check(self._lib.dlp_context_shape_key(self._h, C.byref(k), C.byref(src)))
        return ({name: getattr(k, name) for name, _ in k._fields_},
                ("heuristic", "table", "autotune")[src.value])

    def set_counter_bases(self, stampBase: int, localGen: int):
        """Moves the iteration-stamp base and LDS generation counter up
        (tests of the wrap-around guards)."""
        check(self._lib.dlp_context_set_counter_bases(self._h, stampBase, localGen))

    def counter_bases(self):
        sb, lg = C.c_int(), C.c_int()
        check(self._lib.dlp_context_counter_bases(self._h, C.byref(sb), C.byref(lg)))
        return sb.value, lg.value


class Multi:
    """Many independent graphs, one propagation launch per batch index
    (dlp_multi_*).  `graphs` is a list of Inputs."""

    def __init__(self, graphs, cfgs=None):
        self._lib = load()
        self.graphs = list(graphs)
        n = len(self.graphs)
        self.cfgs = cfgs if cfgs is not None else [make_config(g) for g in self.graphs]
        self._cfgArr = (_lib.Config * n)(*self.cfgs)
        self._gv = (_lib.GraphView * n)(*[g.graph_view() for g in self.graphs])
        self._bv = (_lib.BatchView * n)(*[g.batch_view() for g in self.graphs])
        self._tl = (C.POINTER(C.c_float) * n)(*[_ptr(g.trueLabels, C.c_float)
                                                 for g in self.graphs])
        self._h = C.c_void_p()
        check(self._lib.dlp_multi_create(C.byref(self._h), n, self._cfgArr, self._gv, self._tl,
                                         self._bv))

    def close(self):
        if self._h:
            self._lib.dlp_multi_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, localShape=0, thresholdPercent=0, backoff=0, useGraphs=0,
                   heavyFrontier=0):
        """0 keeps a setting; useGraphs 1 = HIP-graph replay, 2 = enqueue;
        heavyFrontier -1 = graphs never leave the shared launch."""
        t = _lib.MultiTuning(localShape, thresholdPercent, backoff, useGraphs, heavyFrontier)
        check(self._lib.dlp_multi_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = _lib.MultiTuning()
        check(self._lib.dlp_multi_get_tuning(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_}

    def reset(self):
        check(self._lib.dlp_multi_reset(self._h))

    def run_batch(self, b: int):
        check(self._lib.dlp_multi_run_batch(self._h, b))

    def run_all(self):
        check(self._lib.dlp_multi_run_all(self._h))

    def synchronize(self):
        check(self._lib.dlp_multi_synchronize(self._h))

    def labels(self, g: int) -> np.ndarray:
        out = np.empty(self.graphs[g].numNodes, np.float32)
        check(self._lib.dlp_multi_read_labels(self._h, g, _ptr(out, C.c_float)))
        return out

    def stats(self, g: int, b: int) -> dict:
        st = _lib.BatchStats()
        check(self._lib.dlp_multi_batch_stats(self._h, g, b, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def batch_time_ms(self, b: int) -> float:
        ms = C.c_float()
        check(self._lib.dlp_multi_batch_time(self._h, b, C.byref(ms)))
        return ms.value

    def device_bytes(self) -> int:
        v = C.c_int64()
        check(self._lib.dlp_multi_device_bytes(self._h, C.byref(v)))
        return v.value

    def autotune(self, trialBatches: int = 0):
        """Times work-group shapes x hand-back frontiers on the first
        trialBatches batch indices (all when 0); keeps the fastest."""
        t = _lib.MultiTuning()
        ms = C.c_float()
        check(self._lib.dlp_multi_autotune(self._h, trialBatches, C.byref(t), C.byref(ms)))
        return {name: getattr(t, name) for name, _ in t._fields_}, ms.value

    def tuning_source(self) -> str:
        src = C.c_int()
        check(self._lib.dlp_multi_tuning_source(self._h, C.byref(src)))
        return ("defaults", "table", "autotune")[src.value]

    def set_counter_bases(self, g: int, stampBase: int, localGen: int):
        check(self._lib.dlp_multi_set_counter_bases(self._h, g, stampBase, localGen))

    def counter_bases(self, g: int):
        sb, lg = C.c_int(), C.c_int()
        check(self._lib.dlp_multi_counter_bases(self._h, g, C.byref(sb), C.byref(lg)))
        return sb.value, lg.value

    def run_next(self, graphs):
        """The next batch of each listed graph, together."""
        arr = np.ascontiguousarray(graphs, np.int32)
        check(self._lib.dlp_multi_run_next(self._h, _ptr(arr, C.c_int), len(arr)))

    def set_scratch_pool(self, poolBytes: int):
        """poolBytes > 0: batches draw their scratch from one shared pool of
        that size; 0: every graph its own scratch (the default)."""
        check(self._lib.dlp_multi_set_scratch_pool(self._h, int(poolBytes)))

    def pool_stats(self) -> dict:
        st = _lib.PoolStats()
        check(self._lib.dlp_multi_pool_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def batch_scratch_bytes(self, g: int, b: int) -> int:
        v = C.c_int64()
        check(self._lib.dlp_multi_batch_scratch_bytes(self._h, g, b, C.byref(v)))
        return v.value

    def serve(self, requests, maxSlots: int = 64, hbmBudgetBytes: int = 0,
              plain: bool = False, handBackLanes: int = 0, sliceIterations: int = 0,
              sliceMs: float = 0.0, handBackAfter: int = 0):
        """requests: sequence of (graph, arrival seconds), sorted by arrival.
        Returns (completion seconds per request, steps).  sliceIterations
        -1 with sliceMs < 0 runs every step to completion (step-level
        admission).  handBackAfter: with hand-back on, a graph still running
        at this iteration of its batch goes to a hand-back lane (0 = 1024,
        -1 = never)."""
        n = len(requests)
        reqs = (_lib.ServeRequest * max(n, 1))(*[_lib.ServeRequest(int(g), float(t))
                                                  for g, t in requests])
        cfg = _lib.ServeConfig(maxSlots, hbmBudgetBytes, 1 if plain else 0, handBackLanes,
                               sliceIterations, sliceMs, handBackAfter)
        done = np.zeros(max(n, 1), np.float64)
        steps = C.c_int()
        check(self._lib.dlp_multi_serve(self._h, reqs, n, C.byref(cfg),
                                        _ptr(done, C.c_double), C.byref(steps)))
        return done[:n], steps.value


def graph_symmetric(rowPtr, colInd) -> bool:
    """Whether a host CSR names v in u's row exactly as often as u in v's row
    (entries < 0 left out); no GPU needed."""
    rp = np.ascontiguousarray(rowPtr, np.int32)
    ci = np.ascontiguousarray(colInd, np.int32)
    vals = np.zeros(max(len(ci), 1), np.float32)
    g = _lib.GraphView()
    g.numNodes, g.numEdges = len(rp) - 1, len(ci)
    g.rowPtr, g.colInd, g.values = _ptr(rp, C.c_int), _ptr(ci, C.c_int), _ptr(vals, C.c_float)
    out = C.c_int()
    check(load().dlp_graph_symmetric(C.byref(g), C.byref(out)))
    return bool(out.value)


def autotune_heuristic(inputs: Inputs):
    """The degree-histogram heuristic for a graph: (shape key, tuning)."""
    k, t = _lib.ShapeKey(), _lib.Tuning()
    g = inputs.graph_view()
    check(load().dlp_autotune_heuristic(C.byref(g), C.byref(k), C.byref(t)))
    return ({name: getattr(k, name) for name, _ in k._fields_},
            {name: getattr(t, name) for name, _ in t._fields_})


def tuning_table_load(path: str) -> int:
    n = C.c_int()
    check(load().dlp_tuning_table_load(path.encode(), C.byref(n)))
    return n.value


def tuning_table_save(path: str) -> int:
    n = C.c_int()
    check(load().dlp_tuning_table_save(path.encode(), C.byref(n)))
    return n.value


def tuning_table_clear() -> None:
    check(load().dlp_tuning_table_clear())


def counter_guard_plan(stampBound: int, genBound: int, maxIterations: int) -> dict:
    """The library's wrap-around guard decision before one batch (host
    arithmetic only)."""
    lib = load()
    rb, rg = C.c_int(), C.c_int()
    sa, ga = C.c_int64(), C.c_int64()
    check(lib.dlp_counter_guard_plan(stampBound, genBound, maxIterations, C.byref(rb),
                                     C.byref(rg), C.byref(sa), C.byref(ga)))
    return {"rebaseStamps": bool(rb.value), "restartGens": bool(rg.value),
            "stampBound": sa.value, "genBound": ga.value}


def batched_workload(numGraphs: int = 256, numBatches: int = 10, seed: int = 42,
                     minNodes: int = 2000, maxNodes: int = 200000, degrees=(3, 5, 7)):
    """The batched workload: numGraphs graphs with vertex counts
    log-uniform in [minNodes, maxNodes] and average degree drawn from
    `degrees`, both from numpy.random.default_rng(seed); graph g is generated
    with the reference generator and seed + g."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(minNodes), np.log(maxNodes), numGraphs)))
    degs = rng.choice(np.asarray(degrees), numGraphs)
    return [generate(int(n), int(d), numBatches, seed + g)
            for g, (n, d) in enumerate(zip(sizes, degs))]
