"""The allocation-free left-right check (hbp_lr_check_fill_dev_noalloc) against the allocating
entry point it mirrors (hbp_lr_check_fill_dev), through the C ABI on device buffers: two calls
in a row, the second on a smaller image, with and without caller masks; the context's memory
pool stays untouched; images above the scratch's size are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from hbp_amd import lib as H

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class Dev:
    """A device buffer holding a copy of a host array (hipMalloc'ed by the test itself)."""

    def __init__(self, a):
        self.shape, self.dtype = a.shape, a.dtype
        self.nbytes = a.nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(self.nbytes, 1))) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def _inputs(seed, W, Hh, D):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:Hh, 0:W].astype(np.float32)
    # piecewise-planar disparities with noise, so that some pixels fail and fills happen
    dl = np.clip(0.02 * x + 0.01 * y + rng.integers(0, D // 2, (Hh, 1)), 0, D - 1).astype(np.float32)
    dr = dl.copy()
    bad = rng.random((Hh, W)) < 0.15
    dr[bad] = rng.uniform(0, D - 1, bad.sum()).astype(np.float32)
    dl[rng.random((Hh, W)) < 0.05] = np.float32(rng.uniform(0, D - 1))
    plane = rng.normal(0, 1, (Hh, W, 3)).astype(np.float32)
    return dl, dr, plane


def _run(ctx, fn, W, Hh, dl, dr, plane, masks):
    bufs = [Dev(dl), Dev(dr), Dev(plane)]
    mk = [Dev(np.full((Hh, W), 7, np.uint8)) for _ in range(3)] if masks else []
    ptrs = [m.p for m in mk] if masks else [None, None, None]
    rc = fn(ctx.h, bufs[0].p, bufs[1].p, bufs[2].p, W, Hh, *ptrs)
    assert rc == 0, ctx.lib.hbp_last_error()
    out = [bufs[2].get()] + [m.get() for m in mk]
    for b in bufs + mk:
        b.free()
    return out


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                                 np.ascontiguousarray(b).view(np.uint8))


def test_noalloc_equals_the_allocating_entry_point_on_two_calls():
    ctx = H.Ctx(0)
    try:
        lib = ctx.lib
        pool0 = lib.hbp_ctx_pool_bytes(ctx.h)
        for masks in (True, False):
            for seed, W, Hh, D in ((1, 1280, 720, 128), (2, 333, 201, 64)):   # second call smaller
                print(f"[case] noalloc lr W={W} H={Hh} masks={masks}", flush=True)
                dl, dr, plane = _inputs(seed, W, Hh, D)
                got = _run(ctx, lib.hbp_lr_check_fill_dev_noalloc, W, Hh, dl, dr, plane, masks)
                assert lib.hbp_ctx_pool_bytes(ctx.h) == pool0 == 0     # nothing allocated
                ref = _run(ctx, lib.hbp_lr_check_fill_dev, W, Hh, dl, dr, plane, masks)
                for a, b in zip(got, ref):
                    assert same_bits(a, b)
                assert not same_bits(got[0], plane)                    # some pixels were filled
                pool0 = 0
                # the allocating call fills the pool; start the next comparison from a fresh one
                ctx.close()
                ctx = H.Ctx(0)
                lib = ctx.lib
    finally:
        ctx.close()


def test_noalloc_refuses_images_above_its_scratch():
    ctx = H.Ctx(0)
    try:
        W, Hh = 4096, 2049            # 2^23 + 4096 pixels: no buffers are read, nothing launches
        one = Dev(np.zeros(4, np.float32))
        rc = ctx.lib.hbp_lr_check_fill_dev_noalloc(ctx.h, one.p, one.p, one.p, W, Hh, None, None, None)
        one.free()
        assert rc != 0 and b"scratch" in ctx.lib.hbp_last_error()
    finally:
        ctx.close()
