"""The allocation-free sort (b3h_radix_sort32_noalloc) against the allocating
one (b3h_radix_sort32) through the C ABI: two calls in a row on the shared
device scratch, the second smaller, each bit-exact with the allocating sort and
with numpy's stable sort; sizes above the scratch are refused."""
import ctypes

import numpy as np
import pytest

from bullet3_amd import capi

pytestmark = pytest.mark.gpu


def _keys(rng, n, bits):
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, min(hi, max(2, n // 4)), n, dtype=np.uint64).astype(np.uint32)
    keys ^= rng.integers(0, 2, n, dtype=np.uint64).astype(np.uint32) << np.uint32(bits - 1)
    return keys


@pytest.mark.parametrize("bits", [13, 32])
def test_noalloc_sort_matches_allocating_sort(bits):
    rng = np.random.default_rng(bits)
    # the full scratch first, then a smaller ragged input that leaves stale
    # entries of the first call behind in the shared scratch
    for n in (1 << 20, 100003, 257):
        keys = _keys(rng, n, bits)
        vals = rng.permutation(n).astype(np.uint32)
        k0, v0 = capi.radix_sort32(keys, vals, bits)
        k1, v1 = capi.radix_sort32(keys, vals, bits, entry="b3h_radix_sort32_noalloc")
        assert np.array_equal(k1, k0) and np.array_equal(v1, v0), n
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k1, keys[order]) and np.array_equal(v1, vals[order]), n


def test_noalloc_sort_refuses_oversize_input():
    n = (1 << 20) + 1
    buf = capi.DeviceBuffer(4)
    r = capi.lib().b3h_radix_sort32_noalloc(ctypes.c_void_p(buf.ptr), ctypes.c_void_p(buf.ptr), n, 32)
    assert r == -2  # B3H_ERR_CAPACITY, checked before any device access
    assert b"exceed" in capi.lib().b3h_last_error()
    assert capi.lib().b3h_radix_sort32_noalloc(None, None, 0, 32) == 0
