"""The allocation-free prefix scan (deme_prefix_scan_u32_noalloc) against the allocating one (deme_prefix_scan_u32),
through the C ABI: bit-exact outputs and totals on two calls in a row, the second smaller (the fixed workspace keeps
the first call's tile sums, which the second must not read), then on the largest size it takes, and a clear error
beyond that size (rejected on the host before any launch).
"""
import ctypes as C

import numpy as np
import pytest

from deme_amd.abi import DEME_OK, load

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 28


def _scan_pair(lib, ctx, x):
    """Both entry points on the same input; returns (noalloc out, noalloc total, allocating out, allocating total)."""
    from devmem import DeviceArray

    n = len(x)
    din, out_a, out_b, tot_a = DeviceArray.of(x), DeviceArray(4 * (n + 1)), DeviceArray(4 * (n + 1)), DeviceArray(8)
    try:
        st = lib.deme_prefix_scan_u32_noalloc(ctx, din.ptr if n else None, out_a.ptr, n, tot_a.ptr)
        assert st == DEME_OK, lib.deme_last_error(ctx)
        assert lib.deme_sync(ctx) == DEME_OK
        tot_b = C.c_uint64(0)
        st = lib.deme_prefix_scan_u32(ctx, din.ptr if n else None, out_b.ptr, n, C.byref(tot_b))
        assert st == DEME_OK, lib.deme_last_error(ctx)
        return (out_a.get(np.uint32, n + 1), int(tot_a.get(np.uint64, 1)[0]), out_b.get(np.uint32, n + 1),
                tot_b.value)
    finally:
        for d in (din, out_a, out_b, tot_a):
            d.close()


def _check(lib, ctx, x):
    a, ta, b, tb = _scan_pair(lib, ctx, x)
    ref = np.zeros(len(x) + 1, np.uint64)
    ref[1:] = np.cumsum(x, dtype=np.uint64)
    assert ta == tb == int(ref[-1])
    assert np.array_equal(a, b)
    assert np.array_equal(a, (ref & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def test_noalloc_scan_equals_allocating_scan_on_consecutive_calls():
    lib = load()
    h = C.c_void_p()
    assert lib.deme_ctx_create(0, None, C.byref(h)) == DEME_OK
    rng = np.random.default_rng(11)
    try:
        # first call on the full-tile path with a recursive scan of the tile sums (beyond 4.2M elements), the second
        # smaller on the folded half-tile path, then a single tile and an empty scan
        for n in (4096 * 2048 + 1, 100_003, 4000, 0):
            print(f"noalloc scan n={n}", flush=True)
            x = rng.integers(0, 1 << 16, n, dtype=np.uint32) if n else np.zeros(0, np.uint32)
            if n > 3:
                x[rng.integers(0, n, 3)] = 0xFFFFFFFF  # totals beyond 32 bits
            _check(lib, h, x)
    finally:
        lib.deme_ctx_destroy(h)


def test_noalloc_scan_largest_size_then_smaller_and_rejects_beyond():
    from devmem import DeviceArray

    lib = load()
    h = C.c_void_p()
    assert lib.deme_ctx_create(0, None, C.byref(h)) == DEME_OK
    rng = np.random.default_rng(12)
    try:
        for n in (NOALLOC_MAX, 12_345):
            print(f"noalloc scan n={n}", flush=True)
            x = rng.integers(0, 1 << 8, n, dtype=np.uint32)
            _check(lib, h, x)
        # one element more than the workspace covers: a host-side error, nothing launched
        small = DeviceArray(16)
        try:
            st = lib.deme_prefix_scan_u32_noalloc(h, small.ptr, small.ptr, NOALLOC_MAX + 1, None)
        finally:
            small.close()
        assert st != DEME_OK
        assert b"2^28" in lib.deme_last_error(h)
    finally:
        lib.deme_ctx_destroy(h)
