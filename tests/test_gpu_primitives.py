"""The project's device primitives that replace the reference's Thrust sorts and scans
(segment.cu:1058-1085, CRTrees.cu:395): the stable LSD radix sort (one-workgroup and
multi-tile forms, 32- and 64-bit keys, with and without values, any end bit) against numpy's
stable argsort, and the exclusive scan (one tile and many) against numpy's cumulative sum,
through hbp_primitive_sort / hbp_primitive_scan."""
import ctypes as C

import numpy as np
import pytest

from hbp_amd import lib as H

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def sort_dev(ctx, keys, vals, end_bit):
    k = np.ascontiguousarray(keys).copy()
    v = None if vals is None else np.ascontiguousarray(vals, np.uint32).copy()
    H._check(ctx.lib.hbp_primitive_sort(ctx.h, k.ctypes.data_as(C.c_void_p), k.dtype.itemsize,
                                        None if v is None else v.ctypes.data_as(C.c_void_p), len(k), end_bit))
    return k, v


@pytest.mark.parametrize("n", [0, 1, 2, 63, 4095, 4096, 4097, 65537, 1 << 20, 2_900_000, 8_388_609])
@pytest.mark.parametrize("key_bits", [32, 64])
def test_radix_sort_is_numpys_stable_sort(ctx, n, key_bits):
    """Up to 2048 tiles of 4096 keys a pass scans each digit's row of tile counts in one launch;
    above (8388609 keys) the whole digit-major table is scanned in tiles."""
    rng = np.random.default_rng(n * 7 + key_bits)
    for end_bit in sorted({1, 7, 8, 13, 24, key_bits}):
        say(f"sort n={n} keys={key_bits} end_bit={end_bit}")
        dt = np.uint32 if key_bits == 32 else np.uint64
        hi = 1 << end_bit
        # few distinct values as well as many, so that stability is exercised
        pool = rng.integers(0, hi, size=max(1, min(n, 50)), dtype=np.uint64, endpoint=False)
        keys = (pool[rng.integers(0, len(pool), n)] if n and end_bit > 6 else
                rng.integers(0, hi, n, dtype=np.uint64)).astype(dt)
        vals = np.arange(n, dtype=np.uint32)
        order = np.argsort(keys, kind="stable")
        k, v = sort_dev(ctx, keys, vals, end_bit)
        assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
        if key_bits == 64:
            k2, _ = sort_dev(ctx, keys, None, end_bit)
            assert np.array_equal(k2, keys[order])


def test_radix_sort_refuses_keys_above_end_bit(ctx):
    keys = np.array([1, 2, 1 << 9], np.uint32)
    with pytest.raises(H.HbpError, match="end_bit"):
        sort_dev(ctx, keys, np.zeros(3, np.uint32), 9)


@pytest.mark.parametrize("n", [0, 1, 1000, 4096, 4097, 123457, 3_000_001])
def test_exclusive_scan_is_the_cumulative_sum(ctx, n):
    say(f"scan n={n}")
    rng = np.random.default_rng(n)
    x = rng.integers(0, 600, n).astype(np.int32)
    out = np.full(n, -7, np.int32)
    H._check(ctx.lib.hbp_primitive_scan(ctx.h, x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), n))
    ref = np.concatenate([[0], np.cumsum(x, dtype=np.int64)[:-1]]).astype(np.int32) if n else x
    assert np.array_equal(out, ref)
