"""ctypes binding of libdynlp.so (the C-ABI declared in include/dynlp/dynlp_c.h).

The shared library is built in-tree (``labelpropcuda_amd/libdynlp.so``) by
``__graft_entry__.build()`` / ``make -C labelpropcuda_amd/csrc``.  Loading
fails loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdynlp.so")
CLI_PATH = os.path.join(_HERE, "bin", "dynLP")

DLP_OK = 0
DLP_ERR_INVALID = 1
DLP_ERR_HIP = 2
DLP_ERR_NOMEM = 3
DLP_ERR_STATE = 4
DLP_ERR_NODEVICE = 5
DLP_ERR_INTERNAL = 6

VERTEX_ACTIVE = 0
VERTEX_DELETED = 1
VERTEX_GROUND_TRUTH = 2

i32p = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)


class Config(C.Structure):
    _fields_ = [
        ("totalNodes", C.c_int),
        ("avgDegree", C.c_int),
        ("numBatches", C.c_int),
        ("seed", C.c_uint),
        ("tau", C.c_float),
        ("delta", C.c_float),
        ("maxIterations", C.c_int),
        ("groundTruthFraction", C.c_float),
        ("unlabeledFraction", C.c_float),
        ("deletedFraction", C.c_float),
        ("validate", C.c_int),
        ("verbose", C.c_int),
    ]


class GraphView(C.Structure):
    _fields_ = [
        ("numNodes", C.c_int),
        ("numEdges", C.c_int),
        ("rowPtr", i32p),
        ("colInd", i32p),
        ("values", f32p),
    ]


class BatchView(C.Structure):
    _fields_ = [
        ("numBatches", C.c_int),
        ("insertedOffsets", i64p),
        ("insertedNodeIds", i32p),
        ("groundTruthOffsets", i64p),
        ("groundTruthNodeIds", i32p),
        ("groundTruthLabels", f32p),
        ("deletedOffsets", i64p),
        ("deletedNodeIds", i32p),
        ("edgeOffsets", i64p),
        ("newEdgeSrc", i32p),
        ("newEdgeDst", i32p),
        ("newEdgeWeight", f32p),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("batch", C.c_int),
        ("numInserted", C.c_int),
        ("numDeleted", C.c_int),
        ("numNewEdges", C.c_int),
        ("numSubEdges", C.c_int),
        ("numComponents", C.c_int),
        ("numMarked", C.c_int),
        ("frontier0", C.c_int),
        ("iterations", C.c_int),
        ("numTouched", C.c_int),
        ("updates", C.c_int64),
        ("timeMs", C.c_float),
        ("accuracy", C.c_float),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("propGroupLanes", C.c_int),
        ("propBlocks", C.c_int),
        ("chunkIterations", C.c_int),
        ("localThreshold", C.c_int),
        ("localLdsBytes", C.c_int),
        ("localShape", C.c_int),
        ("localOwnerAfter", C.c_int),
        ("waveRowMin", C.c_int),
        ("blockRowMin", C.c_int),
        ("maskOwnerMax", C.c_int),
    ]


class Trace(C.Structure):
    _fields_ = [
        ("csrRowPtr", i32p),
        ("csrColInd", i32p),
        ("csrValues", f32p),
        ("affected_N", i32p),
        ("subSrc", i32p),
        ("subDst", i32p),
        ("subW", f32p),
        ("componentId_N", i32p),
        ("initLabels_N", f32p),
        ("frontierSizes", i32p),
        ("maxFrontierSizes", C.c_int),
        ("numSubEdges", C.c_int),
        ("numComponents", C.c_int),
        ("numFrontierSizes", C.c_int),
    ]


class MultiTuning(C.Structure):
    _fields_ = [
        ("localShape", C.c_int),
        ("thresholdPercent", C.c_int),
        ("backoff", C.c_int),
        ("useGraphs", C.c_int),
        ("heavyFrontier", C.c_int),
    ]


class ServeRequest(C.Structure):
    _fields_ = [("graph", C.c_int), ("arrival", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("maxSlots", C.c_int), ("hbmBudgetBytes", C.c_int64), ("plain", C.c_int),
                ("handBackLanes", C.c_int), ("sliceIterations", C.c_int),
                ("sliceMs", C.c_double), ("handBackAfter", C.c_int)]


class PoolStats(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("inUse", C.c_int64), ("peak", C.c_int64),
                ("leases", C.c_int64), ("deferrals", C.c_int64)]


class PropProfile(C.Structure):
    _fields_ = [
        ("launches", C.c_int64),
        ("kernelMs", C.c_double),
        ("edges", C.c_int64),
        ("changedEdges", C.c_int64),
        ("appended", C.c_int64),
        ("localLaunches", C.c_int64),
        ("localKernelMs", C.c_double),
        ("localUpdates", C.c_int64),
        ("localEdges", C.c_int64),
        ("localChangedEdges", C.c_int64),
        ("localAppended", C.c_int64),
        ("localOwnerUpdates", C.c_int64),
        ("localBuilds", C.c_int64),
        ("localWaveUpdates", C.c_int64),
    ]


class IrlpTuning(C.Structure):
    _fields_ = [
        ("groupLanes", C.c_int),
        ("blocks", C.c_int),
        ("firstChunk", C.c_int),
        ("maxChunk", C.c_int),
    ]


class IrlpResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("skipped", C.c_int),
        ("activeEdges", C.c_int),
        ("timeMs", C.c_float),
        ("loopMs", C.c_float),
        ("accuracy", C.c_float),
    ]


class ShapeKey(C.Structure):
    _fields_ = [
        ("logNodes", C.c_int),
        ("meanDegree", C.c_int),
        ("logMaxDegree", C.c_int),
    ]


# name -> (restype, argtypes)
_VP = C.c_void_p
SIGNATURES = {
    "dlp_last_error": (C.c_char_p, []),
    "dlp_version": (C.c_char_p, []),
    "dlp_device_count": (C.c_int, [i32p]),
    "dlp_mark_deletions": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.c_int, C.c_int, _VP]),
    "dlp_build_sparsified_subgraph": (
        C.c_int,
        [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_int, _VP, C.c_int, _VP, C.c_float,
         _VP, _VP, _VP, _VP, i32p, i32p, _VP],
    ),
    "dlp_connected_components": (
        C.c_int, [_VP, _VP, C.c_int, _VP, C.c_int, _VP, i32p, C.c_int, _VP]),
    "dlp_label_initialization": (
        C.c_int,
        [_VP, _VP, _VP, _VP, _VP, _VP, C.c_int, _VP, C.c_int, _VP, C.c_int, _VP]),
    "dlp_iterative_propagation": (
        C.c_int,
        [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_int, C.c_float, C.c_int, i32p, _VP]),
    "dlp_iterative_propagation_noalloc": (
        C.c_int,
        [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_int, C.c_float, C.c_int, i32p, _VP]),
    "dlp_config_default": (None, [C.POINTER(Config)]),
    "dlp_run_dynlp": (
        C.c_int,
        [C.POINTER(Config), C.POINTER(GraphView), f32p, C.POINTER(BatchView), f32p, f32p, f32p]),
    "dlp_context_create": (
        C.c_int,
        [C.POINTER(_VP), C.POINTER(Config), C.POINTER(GraphView), f32p, C.POINTER(BatchView),
         _VP]),
    "dlp_context_destroy": (C.c_int, [_VP]),
    "dlp_context_set_tuning": (C.c_int, [_VP, C.POINTER(Tuning)]),
    "dlp_context_get_tuning": (C.c_int, [_VP, C.POINTER(Tuning)]),
    "dlp_context_tau": (C.c_int, [_VP, f32p]),
    "dlp_context_device_bytes": (C.c_int, [_VP, i64p]),
    "dlp_context_reset": (C.c_int, [_VP]),
    "dlp_context_run_batch": (C.c_int, [_VP, C.c_int]),
    "dlp_context_run_all": (C.c_int, [_VP]),
    "dlp_context_synchronize": (C.c_int, [_VP]),
    "dlp_context_read_labels": (C.c_int, [_VP, f32p]),
    "dlp_context_batch_stats": (C.c_int, [_VP, C.c_int, C.POINTER(BatchStats)]),
    "dlp_context_set_trace": (C.c_int, [_VP, C.POINTER(Trace)]),
    "dlp_context_set_profiling": (C.c_int, [_VP, C.c_int]),
    "dlp_context_prop_profile": (C.c_int, [_VP, C.POINTER(PropProfile)]),
    "dlp_inputs_generate": (
        C.c_int, [C.POINTER(_VP), C.c_int, C.c_int, C.c_int, C.c_uint, C.c_float, C.c_float]),
    "dlp_inputs_graph": (C.c_int, [_VP, C.POINTER(GraphView), C.POINTER(f32p)]),
    "dlp_inputs_batches": (C.c_int, [_VP, C.POINTER(BatchView)]),
    "dlp_inputs_free": (None, [_VP]),
    "dlp_multi_create": (C.c_int, [C.POINTER(_VP), C.c_int, C.POINTER(Config),
                                   C.POINTER(GraphView), C.POINTER(C.POINTER(C.c_float)),
                                   C.POINTER(BatchView)]),
    "dlp_multi_destroy": (C.c_int, [_VP]),
    "dlp_multi_set_tuning": (C.c_int, [_VP, C.POINTER(MultiTuning)]),
    "dlp_multi_get_tuning": (C.c_int, [_VP, C.POINTER(MultiTuning)]),
    "dlp_multi_reset": (C.c_int, [_VP]),
    "dlp_multi_run_batch": (C.c_int, [_VP, C.c_int]),
    "dlp_multi_run_all": (C.c_int, [_VP]),
    "dlp_multi_synchronize": (C.c_int, [_VP]),
    "dlp_multi_read_labels": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_float)]),
    "dlp_multi_batch_stats": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(BatchStats)]),
    "dlp_multi_batch_time": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_float)]),
    "dlp_multi_device_bytes": (C.c_int, [_VP, C.POINTER(C.c_int64)]),
    "dlp_multi_run_next": (C.c_int, [_VP, C.POINTER(C.c_int), C.c_int]),
    "dlp_multi_set_scratch_pool": (C.c_int, [_VP, C.c_int64]),
    "dlp_context_row_update": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int)]),
    "dlp_graph_symmetric": (C.c_int, [C.POINTER(GraphView), C.POINTER(C.c_int)]),
    "dlp_multi_pool_stats": (C.c_int, [_VP, C.POINTER(PoolStats)]),
    "dlp_multi_batch_scratch_bytes": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    "dlp_multi_serve": (C.c_int, [_VP, C.POINTER(ServeRequest), C.c_int, C.POINTER(ServeConfig),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "dlp_inputs_generate_stream": (C.c_int, [C.POINTER(_VP), C.c_int, C.c_int,
                                             C.POINTER(C.c_int), C.c_int, C.c_uint, C.c_float,
                                             C.c_float]),
    "dlp_copy_device": (C.c_int, [_VP, _VP, C.c_size_t]),
    "dlp_random_access_rates": (C.c_int, [C.c_int64, C.c_int64, C.POINTER(C.c_double)]),
    "dlp_grid_sync_cost": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_int)]),
    "dlp_write_graph_files": (C.c_int, [C.POINTER(GraphView), f32p, C.c_char_p]),
    "dlp_inputs_from_graph_files": (C.c_int, [C.POINTER(_VP), C.c_char_p, C.c_int, C.c_uint,
                                              C.c_float, C.c_float]),
    "dlp_irlp": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, C.c_int, C.c_float, C.c_int, i32p, _VP]),
    "dlp_irlp_tuning_default": (None, [C.POINTER(IrlpTuning)]),
    "dlp_irlp_ex": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, C.c_int, C.c_float, C.c_int, i32p, _VP,
                              C.POINTER(IrlpTuning), f32p]),
    "dlp_irlp_validation": (C.c_int, [C.POINTER(GraphView), i32p, i32p, f32p, C.c_int, C.c_float,
                                      C.c_int, i32p, i32p]),
    "dlp_irlp_validation_ex": (C.c_int, [C.POINTER(GraphView), i32p, i32p, f32p, C.c_int,
                                         C.c_float, C.c_int, i32p, i32p, C.POINTER(IrlpTuning),
                                         f32p]),
    "dlp_irlp_baseline": (C.c_int, [C.POINTER(Config), C.POINTER(GraphView), f32p,
                                    C.POINTER(BatchView), C.POINTER(IrlpTuning), f32p,
                                    C.POINTER(IrlpResult)]),
    "dlp_device_malloc": (C.c_int, [C.POINTER(_VP), C.c_size_t]),
    "dlp_device_free": (C.c_int, [_VP]),
    "dlp_copy_to_device": (C.c_int, [_VP, _VP, C.c_size_t]),
    "dlp_copy_to_host": (C.c_int, [_VP, _VP, C.c_size_t]),
    "dlp_autotune_heuristic": (C.c_int, [C.POINTER(GraphView), C.POINTER(ShapeKey),
                                         C.POINTER(Tuning)]),
    "dlp_context_autotune": (C.c_int, [_VP, C.c_int, C.POINTER(Tuning), f32p]),
    "dlp_context_shape_key": (C.c_int, [_VP, C.POINTER(ShapeKey), i32p]),
    "dlp_multi_autotune": (C.c_int, [_VP, C.c_int, C.POINTER(MultiTuning), f32p]),
    "dlp_multi_tuning_source": (C.c_int, [_VP, i32p]),
    "dlp_tuning_table_load": (C.c_int, [C.c_char_p, i32p]),
    "dlp_tuning_table_save": (C.c_int, [C.c_char_p, i32p]),
    "dlp_tuning_table_clear": (C.c_int, []),
    "dlp_counter_guard_plan": (C.c_int, [C.c_int64, C.c_int64, C.c_int, i32p, i32p, i64p, i64p]),
    "dlp_context_set_counter_bases": (C.c_int, [_VP, C.c_int, C.c_int]),
    "dlp_context_counter_bases": (C.c_int, [_VP, i32p, i32p]),
    "dlp_multi_set_counter_bases": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int]),
    "dlp_multi_counter_bases": (C.c_int, [_VP, C.c_int, i32p, i32p]),
}

_lib = None


class DynLPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libdynlp error {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load libdynlp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(make -C labelpropcuda_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != DLP_OK:
        msg = load().dlp_last_error().decode(errors="replace")
        raise DynLPError(rc, msg)


def device_count() -> int:
    n = C.c_int(0)
    check(load().dlp_device_count(C.byref(n)))
    return n.value


class DeviceArray:
    """A device copy of a numpy array (dlp_device_malloc / dlp_copy_*), for
    calling the raw-pointer step API."""

    def __init__(self, host):
        import numpy as np
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        lib = load()
        check(lib.dlp_device_malloc(C.byref(self.ptr), max(self.host.nbytes, 1)))
        check(lib.dlp_copy_to_device(self.ptr, self.host.ctypes.data_as(C.c_void_p),
                                     self.host.nbytes))

    def numpy(self):
        out = self.host.copy()
        check(load().dlp_copy_to_host(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            load().dlp_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
