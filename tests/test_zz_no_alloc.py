"""erod_check_finite_noalloc (flags in the library's device scratch array)
against erod_check_finite through the C ABI: two calls in a row, the second
on a smaller batch, each with its own pattern of non-finite rods, so a flag
left in the shared array by the first call would show in the second."""
import ctypes

import numpy as np
import pytest

from elasticapp_amd import load_library
from tests.gpu_helpers import device_block, random_workload

pytestmark = pytest.mark.gpu


def _flags(blk, name):
    L = load_library()
    out = np.full(blk.n_rods, 7, dtype=np.uint8)
    rc = getattr(L, name)(blk._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert rc == 0, L.erod_last_error().decode()
    return out


def _poison(blk, rods, tag):
    for r in rods:
        a = blk.get_rod(tag, r)
        a[a.shape[0] - 1, a.shape[1] // 2] = np.nan if r % 2 else np.inf
        blk.set_rod(tag, r, a)


def test_noalloc_check_equals_the_allocating_one():
    first = device_block(random_workload(40, 2, 120, seed=81, steps=0))
    first.step_fused(20, 0)
    _poison(first, [3, 7, 31, 38], "Position")
    _poison(first, [12], "Director")
    _poison(first, [20], "AngularVelocity")
    a, b = _flags(first, "erod_check_finite"), _flags(first, "erod_check_finite_noalloc")
    assert np.array_equal(a, b) and a.sum() == 6 and set(np.flatnonzero(a)) == {3, 7, 12, 20, 31, 38}
    # a smaller batch next: only its own bad rods are flagged
    second = device_block(random_workload(9, 2, 60, seed=82, steps=0))
    _poison(second, [1], "Velocity")
    a, b = _flags(second, "erod_check_finite"), _flags(second, "erod_check_finite_noalloc")
    assert np.array_equal(a, b) and list(np.flatnonzero(b)) == [1]
    # and a clean batch has no flags
    third = device_block(random_workload(5, 2, 30, seed=83, steps=0))
    assert not _flags(third, "erod_check_finite_noalloc").any()
    for blk in (first, second, third):
        blk.close()


def test_noalloc_check_refuses_more_rods_than_its_array():
    """A batch of 2^20 + 1 two-element rods (about 3.3 GB): the allocating
    check runs, the one without allocation refuses with its limit."""
    L = load_library()
    n = (1 << 20) + 1
    n_elems = np.full(n, 2, dtype=np.uint64)
    h = ctypes.c_void_p()
    assert L.erod_batch_create(n_elems.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n, 1, ctypes.byref(h)) == 0
    try:
        out = np.zeros(n, dtype=np.uint8)
        p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        assert L.erod_check_finite(h, p) == 0 and not out.any()
        assert L.erod_check_finite_noalloc(h, p) == -1
        assert "1,048,576" in L.erod_last_error().decode()
    finally:
        L.erod_batch_destroy(h)
