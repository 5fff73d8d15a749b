"""slak_scan_exclusive_noalloc: the scan with its flags word and partials in
the library's static device scratch (no allocation). Compared with
slak_scan_exclusive on two calls in a row, the second with a smaller input
(the scratch keeps the first call's contents), and with the host cumsum."""
import ctypes

import numpy as np
import pytest

from slak_amd import capi
from slak_amd.mesh import DeviceBuffer, to_host

pytestmark = pytest.mark.gpu

MAX_ITEMS = 268_367_872


def pool_state():
    a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    capi.check(capi.lib().slak_pool_stats(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value


def scan_both(x):
    n = len(x)
    din = DeviceBuffer.from_array(x if n else np.zeros(1, np.int32))
    ref_out = DeviceBuffer(4 * (n + 1))
    out = DeviceBuffer(4 * (n + 1))
    capi.check(capi.lib().slak_scan_exclusive(din.ptr, ref_out.ptr, n, None))
    before = pool_state()
    capi.check(capi.lib().slak_scan_exclusive_noalloc(din.ptr, out.ptr, n, None))
    assert pool_state() == before, "the no-alloc scan touched the device pool"
    got = to_host(out.ptr, n + 1, np.int32)
    ref = to_host(ref_out.ptr, n + 1, np.int32)
    return got, ref


def test_noalloc_scan_equals_scan_twice_second_smaller():
    rng = np.random.default_rng(7)
    for n in (3_000_001, 40_000):  # three-kernel path, then a smaller input
        x = rng.integers(0, 9, size=n).astype(np.int32)
        got, ref = scan_both(x)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, np.concatenate([[0], np.cumsum(x, dtype=np.int64)]).astype(np.int32))


def test_noalloc_scan_small_and_empty():
    rng = np.random.default_rng(8)
    for n in (5000, 0, 1):
        x = rng.integers(0, 100, size=n).astype(np.int32)
        got, ref = scan_both(x)
        np.testing.assert_array_equal(got, ref)


def test_noalloc_scan_rejects_beyond_scratch():
    d = DeviceBuffer(16)
    code = capi.lib().slak_scan_exclusive_noalloc(d.ptr, d.ptr, MAX_ITEMS + 1, None)
    assert code == 1
    assert "scratch capacity" in capi.lib().slak_last_error().decode()


def test_noalloc_scan_overflow_flag_cleared_between_calls():
    """A call whose total overflows int32 raises; the next call starts from a
    cleared flags word and succeeds."""
    big = np.full(50_000, 60_000, np.int32)  # total 3e9 > int32
    din = DeviceBuffer.from_array(big)
    out = DeviceBuffer(4 * (len(big) + 1))
    code = capi.lib().slak_scan_exclusive_noalloc(din.ptr, out.ptr, len(big), None)
    assert code == 1 and "int32" in capi.lib().slak_last_error().decode()
    x = np.arange(1000, dtype=np.int32)
    got, ref = scan_both(x)
    np.testing.assert_array_equal(got, ref)
