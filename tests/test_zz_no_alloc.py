"""dlp_iterative_propagation_noalloc (scratch in a static device array, no
allocation) against dlp_iterative_propagation (library workspace) through
the C-ABI: identical labels, buffers, exit affected flags and iteration
counts, on two calls in a row, the second on a smaller graph, and a size
beyond the static limit is refused."""
import ctypes as C

import numpy as np
import pytest

import labelpropcuda_amd as L
from labelpropcuda_amd import _lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def state_after(inputs, batches):
    """CSR, labels entering propagation, affected flags, device status and GT
    classes of batch `batches - 1` (oracle state)."""
    o = Oracle(inputs)
    for b in range(batches):
        o.run_batch(b)
    rp, ci, vv = o.csr()
    d, _, g = o.status()
    return rp, ci, vv, o.init_labels(), o.affected(), d, g


def run(fn, st, cap):
    rp, ci, vv, init, aff, d, g = st
    keep = [_lib.DeviceArray(x) for x in (rp, ci, vv, init, np.zeros_like(init), d, g, aff)]
    t_rp, t_ci, t_vv, t_lab, t_buf, t_d, t_g, t_aff = keep
    it = C.c_int()
    _lib.check(fn(t_rp.ptr, t_ci.ptr, t_vv.ptr, t_lab.ptr, t_buf.ptr, t_d.ptr, t_g.ptr,
                  t_aff.ptr, len(init), C.c_float(1e-4), cap, C.byref(it), None))
    out = (t_lab.numpy(), t_buf.numpy(), t_aff.numpy(), it.value)
    for k in keep:
        k.free()
    return out


def test_noalloc_matches_workspace_entry_point(gpu_available, lib):
    big = state_after(L.generate(60000, 7, 6, 77), 4)
    small = state_after(L.generate(5000, 5, 10, 42), 6)
    for name, st, cap in (("60000-vertex graph", big, 100000), ("5000-vertex graph", small, 100000),
                          ("5000-vertex graph, cap 9", small, 9)):
        ref = run(lib.dlp_iterative_propagation, st, cap)
        got = run(lib.dlp_iterative_propagation_noalloc, st, cap)
        assert got[3] == ref[3], f"{name}: iterations {got[3]} vs {ref[3]}"
        assert np.array_equal(bits(got[0]), bits(ref[0])), f"{name}: labels"
        assert np.array_equal(bits(got[1]), bits(ref[1])), f"{name}: second buffer"
        assert np.array_equal(got[2], ref[2]), f"{name}: affected flags"


def test_noalloc_refuses_sizes_above_its_limit(gpu_available, lib):
    dummy = _lib.DeviceArray(np.zeros(4, np.int32))
    it = C.c_int()
    p = dummy.ptr
    rc = lib.dlp_iterative_propagation_noalloc(p, p, p, p, p, p, p, p, (1 << 19) + 1,
                                               C.c_float(1e-4), 10, C.byref(it), None)
    dummy.free()
    assert rc == _lib.DLP_ERR_INVALID
    assert b"limit" in lib.dlp_last_error()
