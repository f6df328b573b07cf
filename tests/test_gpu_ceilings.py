"""The device probes that the measurements quote beside the propagation
kernels: random 4-byte access rates, and the cost of a device-wide sync
between dependent rounds (a launch boundary against in-kernel barriers).
Both must measure what they claim: finite positive rates, and every round
of the sync probe seeing all writes of the round before."""
import ctypes as C

import pytest

from labelpropcuda_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("blocks", [1, 64, 512])
def test_grid_sync_probe_sees_every_write(gpu_available, blocks):
    """Launch per round, counter barrier and flag barrier each leave every
    word at the round count, and no bounded wait gave up."""
    us, ok = (C.c_double * 3)(), (C.c_int * 3)()
    _lib.check(_lib.load().dlp_grid_sync_cost(blocks, 301, us, ok))
    assert list(ok) == [1, 1, 1]
    assert all(u > 0.0 for u in us)


def test_grid_sync_probe_rejects_bad_arguments(gpu_available):
    us, ok = (C.c_double * 3)(), (C.c_int * 3)()
    lib = _lib.load()
    assert lib.dlp_grid_sync_cost(0, 10, us, ok) != 0
    assert lib.dlp_grid_sync_cost(4, 0, us, ok) != 0
    assert lib.dlp_grid_sync_cost(1 << 20, 10, us, ok) != 0  # cannot all be resident


def test_random_access_rates_are_positive(gpu_available):
    rates = (C.c_double * 3)()
    _lib.check(_lib.load().dlp_random_access_rates(1 << 20, 1 << 20, rates))
    assert all(r > 0.0 for r in rates)
