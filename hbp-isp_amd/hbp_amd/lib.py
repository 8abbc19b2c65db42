"""ctypes bindings to libhbp_hip.so (the C ABI in include/hbp_hip.h).

The library is loaded from hbp-isp_amd/lib/; nothing falls back to another implementation:
a missing library, a missing symbol or a failing call raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhbp_hip.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "hbp_hip.h")

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_pp = C.POINTER(C.c_void_p)


class HbpError(RuntimeError):
    pass


SIGNATURES = {
    "hbp_ctx_create": (C.c_int, [C.c_int, _pp]),
    "hbp_ctx_destroy": (C.c_int, [_vp]),
    "hbp_ctx_sync": (C.c_int, [_vp]),
    "hbp_ctx_stream": (_vp, [_vp]),
    "hbp_last_error": (C.c_char_p, []),
    "hbp_ctx_pool_bytes": (C.c_longlong, [_vp]),
    "hbp_profile_enable": (C.c_int, [_vp, C.c_int]),
    "hbp_profile_reset": (C.c_int, [_vp]),
    "hbp_profile_read": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_double)]),
    "hbp_profile_names": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "hbp_profile_only": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "hbp_ctx_set_tuning": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "hbp_ctx_get_tuning": (C.c_int, [_vp, C.c_char_p, _ip]),
    "hbp_cost_volume": (C.c_int, [_vp, _u8, _u8, C.c_int, C.c_int, C.c_int, C.c_char_p, _pp]),
    "hbp_cost_volume_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_char_p, _pp]),
    "hbp_cost_to_right": (C.c_int, [_vp, _vp]),
    "hbp_volume_info": (C.c_int, [_vp, _ip, _ip, _ip, _fp, _ip]),
    "hbp_volume_dev_wta": (_vp, [_vp]),
    "hbp_volume_dev_f32": (_vp, [_vp, _vp]),
    "hbp_volume_dev_u8": (_vp, [_vp]),
    "hbp_volume_download": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hbp_cost_volume_f32": (C.c_int, [_vp, _f32, C.c_int, C.c_int, C.c_int, C.c_char_p, _pp]),
    "hbp_cost_volume_f32_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_char_p, _pp]),
    "hbp_volume_download_f32": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hbp_volume_free": (C.c_int, [_vp, _vp]),
    "hbp_pyramid_build": (C.c_int, [_vp, _u8, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _pp]),
    "hbp_pyramid_build_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _pp]),
    "hbp_pyramid_build_target": (C.c_int, [_vp, _u8, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                           _pp]),
    "hbp_pyramid_build_target_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                               C.c_int, _pp]),
    "hbp_pyramid_levels": (C.c_int, [_vp]),
    "hbp_pyramid_level_size": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "hbp_pyramid_dev_clus": (_vp, [_vp]),
    "hbp_pyramid_dev_mean": (_vp, [_vp]),
    "hbp_pyramid_dev_bd": (_vp, [_vp]),
    "hbp_pyramid_dev_bdl": (_vp, [_vp]),
    "hbp_pyramid_download_level": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "hbp_pyramid_free": (C.c_int, [_vp, _vp]),
    "hbp_crtrees_get_clus": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "hbp_crtrees_compact": (C.c_int, [_vp, _vp, C.c_int, _ip]),
    "hbp_bp_create": (C.c_int, [_vp, _vp, _vp, C.c_float, C.c_float, C.c_int, C.c_int, _pp]),
    "hbp_bp_step": (C.c_int, [_vp, _vp, C.c_int, _ip]),
    "hbp_bp_total_sweeps": (C.c_int, [_vp]),
    "hbp_bp_level": (C.c_int, [_vp]),
    "hbp_bp_level_size": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "hbp_bp_run": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hbp_bp_run_batch": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "hbp_pairs_prepare_dev": (C.c_int, [_vp, C.c_int, _vp, C.c_float, C.c_float, C.c_int, C.c_int,
                                        C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "hbp_server_create": (C.c_int, [_vp, C.c_int, C.c_longlong, C.c_float, C.c_float, _pp]),
    "hbp_server_destroy": (C.c_int, [_vp]),
    "hbp_submit": (C.c_int, [_vp, _u8, _u8, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_longlong)]),
    "hbp_server_step": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hbp_poll": (C.c_int, [_vp, C.c_longlong, _vp, _vp]),
    "hbp_cancel": (C.c_int, [_vp, C.c_longlong]),
    "hbp_server_set_max_pending": (C.c_int, [_vp, C.c_int]),
    "hbp_server_stats": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "hbp_profile_copy": (C.c_int, [_vp, C.c_longlong, C.c_int]),
    "hbp_primitive_sort": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_longlong, C.c_int]),
    "hbp_primitive_scan": (C.c_int, [_vp, _vp, _vp, C.c_longlong]),
    "hbp_inject_fault": (C.c_int, [C.c_char_p, C.c_longlong]),
    "hbp_bp_dev_disp": (_vp, [_vp]),
    "hbp_bp_dev_plane": (_vp, [_vp]),
    "hbp_bp_get_state": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "hbp_bp_free": (C.c_int, [_vp, _vp]),
    "hbp_schedule_sweeps": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "hbp_weighted_median": (C.c_int, [_vp, _u8, _u8, _f32, C.c_int, C.c_int]),
    "hbp_weighted_median_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int]),
    "hbp_median7": (C.c_int, [_vp, _f32, C.c_int, C.c_int]),
    "hbp_median7_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int]),
    "hbp_eval_bad_pixels": (C.c_int, [_vp, _f32, _f32, _u8, C.c_int, C.c_int, C.c_double,
                                      C.POINTER(C.c_longlong)]),
    "hbp_lr_check_fill": (C.c_int, [_vp, _f32, _f32, _f32, C.c_int, C.c_int, _vp, _vp, _vp]),
    "hbp_lr_check_fill_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp]),
    "hbp_lr_check_fill_dev_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp]),
    "hbp_plane_to_disp": (C.c_int, [_vp, _f32, C.c_int, C.c_int, _f32]),
}

_lib = None


def load():
    """Load libhbp_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HbpError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C hbp-isp_amd)")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise HbpError(load().hbp_last_error().decode())


# Handles are freed explicitly (free()/close()), by their context's close(), or at interpreter
# exit by the atexit hook below, which runs while the HIP runtime is still whole.  Finalisers
# that run after it (or during interpreter shutdown) make no C call at all.
_shutting_down = False
_contexts = weakref.WeakSet()


def _alive():
    return not _shutting_down and not sys.is_finalizing()


@atexit.register
def _close_all():
    global _shutting_down
    for c in list(_contexts):
        try:
            c.close()
        except Exception:
            pass
    _shutting_down = True


def inject_fault(what, after):
    """Host-forced failure of the after-th next device allocation ("alloc") or batched round
    ("round"); 0 disarms (hbp_inject_fault)."""
    _check(load().hbp_inject_fault(what.encode(), int(after)))


class PairDev(C.Structure):
    """hbp_pair_dev: one pair's device images and sizes for hbp_pairs_prepare_dev"""
    _fields_ = [("left", C.c_void_p), ("right", C.c_void_p), ("width", C.c_int), ("height", C.c_int),
                ("labels", C.c_int), ("linkage", C.c_int), ("num_nb", C.c_int), ("min_level", C.c_int)]


class _Handle:
    """A library object owned by a context: freed before the context is."""

    _free_fn = None

    def _adopt(self, ctx, h):
        self.ctx, self.h = ctx, h
        ctx._children.add(self)

    @classmethod
    def _wrap(cls, ctx, h, **attrs):
        """an object for a handle the library made (already owned by ctx)"""
        obj = cls.__new__(cls)
        obj._adopt(ctx, h)
        for k, v in attrs.items():
            setattr(obj, k, v)
        return obj

    def free(self):
        h = getattr(self, "h", None)
        if not h:
            return
        self.h = None
        ctx = self.ctx
        ctx._children.discard(self)
        if ctx.h and _alive():
            getattr(ctx.lib, self._free_fn)(ctx.h, h)

    def __del__(self):
        if _alive():
            self.free()


class Ctx:
    """One device context (one HIP stream, one device memory pool)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        _check(lib.hbp_ctx_create(device, C.byref(h)))
        self.h = h
        self.lib = lib
        self._children = weakref.WeakSet()
        _contexts.add(self)

    def close(self):
        """Frees the context's servers, BP objects, pyramids and volumes, then the context."""
        if not getattr(self, "h", None):
            return
        kids = list(self._children)
        order = {"Server": 0, "BP": 1, "Pyramid": 2, "Volume": 3}
        for k in sorted(kids, key=lambda o: order.get(type(o).__name__, 4)):
            k.free()
        rc = self.lib.hbp_ctx_destroy(self.h)
        self.h = None
        _check(rc)

    def __del__(self):
        if _alive():
            self.close()

    def sync(self):
        _check(self.lib.hbp_ctx_sync(self.h))

    # ---- kernel-form tuning (results are the same for every value)
    TUNING = ("coarse_min_pix", "single_hb", "batch_hb", "hamming_cols", "seg_len", "batch_graph", "fuse",
              "pipe_batch", "pipe_blocks", "pipe_hb", "pipe_window", "run_max_e",
              "crt_rounds", "crt_jumps", "crt_max_jumps")

    def set_tuning(self, name, value):
        _check(self.lib.hbp_ctx_set_tuning(self.h, name.encode(), int(value)))

    def get_tuning(self, name=None):
        if name is None:
            return {n: self.get_tuning(n) for n in self.TUNING}
        v = C.c_int()
        _check(self.lib.hbp_ctx_get_tuning(self.h, name.encode(), C.byref(v)))
        return v.value

    # ---- kernel timing (HIP events on the context stream)
    def profile(self, on=True):
        _check(self.lib.hbp_profile_enable(self.h, int(on)))

    def profile_reset(self):
        _check(self.lib.hbp_profile_reset(self.h))

    def profile_only(self, prefix=None, every=1):
        _check(self.lib.hbp_profile_only(self.h, prefix.encode() if prefix else None, every))

    def profile_stats(self):
        n = self.lib.hbp_profile_names(self.h, None, 0)
        buf = C.create_string_buffer(max(n, 1))
        self.lib.hbp_profile_names(self.h, buf, n)
        out = {}
        for name in filter(None, buf.value.decode().split(",")):
            ms, cnt, by = C.c_double(), C.c_longlong(), C.c_double()
            _check(self.lib.hbp_profile_read(self.h, name.encode(), C.byref(ms), C.byref(cnt), C.byref(by)))
            out[name] = dict(ms=ms.value, launches=cnt.value, bytes=by.value)
        return out

    # ---- (a)
    def cost_volume(self, left, right, D, right_view=False, cost="Census"):
        """compute_stereo_cost(cost) on two BGR images: "Census" (u8 volume) or "AD" (f32)"""
        return Volume(self, left, right, D, right_view, cost=cost)

    def cost_volume_f32(self, vol, right_view=False):
        """the "MC-CNN" cost: a given f32 volume [D][H][W]"""
        return Volume(self, None, None, vol.shape[0], right_view, f32=vol)

    # ---- (b)
    def pyramid(self, bgr, num_nb=4, link=2, target_clus=1, sigma=0.0):
        """SegHAC(bgr, link, num_nb, sigma, LabelISPFormat, target_clus).run_ms(); link 0 Min,
        1 Max, 2 Centroid (default), 3 Ward; sigma > 0 blurs first (round(3 sigma) <= 8)"""
        return Pyramid(self, bgr, num_nb, link, target_clus, sigma)

    # ---- (c)
    def bp(self, pyr, vol, alpha=1000.1, tau_s=0.6, min_level=2, use_cost_vol=1):
        return BP(self, pyr, vol, alpha, tau_s, min_level, use_cost_vol)

    def prepare_pairs_dev(self, pairs, alpha=1000.1, tau_s=0.6, use_cost_vol=1, workers=4, link=2, num_nb=4):
        """Cost volumes, pyramids and BP objects of many pairs built by `workers` host threads on
        streams of their own (hbp_pairs_prepare_dev).  pairs: (left_dev, right_dev, W, H, D,
        min_level) with device pointers to the BGR images; returns [(Volume, Pyramid, BP)]."""
        n = len(pairs)
        arr = (PairDev * max(n, 1))(*[PairDev(l, r, W, H, D, link, num_nb, ml) for l, r, W, H, D, ml in pairs])
        vols, pyrs, bps = (_vp * max(n, 1))(), (_vp * max(n, 1))(), (_vp * max(n, 1))()
        _check(self.lib.hbp_pairs_prepare_dev(self.h, n, C.cast(arr, _vp), alpha, tau_s, int(use_cost_vol),
                                              int(workers), vols, pyrs, bps))
        out = []
        for i, (_, _, W, H, D, _) in enumerate(pairs):
            v = Volume._wrap(self, vols[i], W=W, H=H, D=D, f32=False)
            p = Pyramid._wrap(self, pyrs[i], W=W, H=H)
            b = BP._wrap(self, bps[i], W=W, H=H, pyr=p, vol=v)
            out.append((v, p, b))
        return out

    def bp_run_batch(self, bps):
        """Runs the remaining sweeps of several BP objects together (hbp_bp_run_batch); each
        object's results are then read with its run()."""
        arr = (C.c_void_p * len(bps))(*[b.h for b in bps])
        _check(self.lib.hbp_bp_run_batch(self.h, len(bps), arr))

    # ---- (d)
    def lr_check_fill(self, dispL, dispR, planeL):
        H, W = dispL.shape
        plane = np.ascontiguousarray(planeL, dtype=np.float32).copy()
        fl = np.zeros((H, W), np.uint8)
        fr = np.zeros((H, W), np.uint8)
        f2 = np.zeros((H, W), np.uint8)
        _check(self.lib.hbp_lr_check_fill(self.h, np.ascontiguousarray(dispL, np.float32),
                                          np.ascontiguousarray(dispR, np.float32), plane, W, H,
                                          _np_ptr(fl), _np_ptr(fr), _np_ptr(f2)))
        return plane, fl, fr, f2

    def weighted_median(self, bgr, fail, plane3):
        """PostProcess::median_filter: the failed pixels' planes (a new array)"""
        H, W = fail.shape
        out = np.ascontiguousarray(plane3, dtype=np.float32).copy()
        _check(self.lib.hbp_weighted_median(self.h, np.ascontiguousarray(bgr, np.uint8),
                                            np.ascontiguousarray(fail, np.uint8), out, W, H))
        return out

    def eval_bad_pixels(self, disp, gt, nonocc, threshold):
        """Evaluator::evaluate's counts: (valid, good), (nonocc, good), (occ, good)"""
        H, W = gt.shape
        if np.shape(disp) != (H, W) or np.shape(nonocc) != (H, W):
            raise ValueError("eval_bad_pixels: disp, gt and nonocc must all be [H][W]")
        cnt = (C.c_longlong * 6)()
        _check(self.lib.hbp_eval_bad_pixels(self.h, np.ascontiguousarray(disp, np.float32),
                                            np.ascontiguousarray(gt, np.float32),
                                            np.ascontiguousarray(nonocc, np.uint8), W, H, float(threshold), cnt))
        return [(cnt[0], cnt[1]), (cnt[2], cnt[3]), (cnt[4], cnt[5])]

    def median7(self, disp):
        """the CLI's 7-tap separable median (a new array)"""
        H, W = disp.shape
        out = np.ascontiguousarray(disp, dtype=np.float32).copy()
        _check(self.lib.hbp_median7(self.h, out, W, H))
        return out

    def plane_to_disp(self, plane3):
        H, W = plane3.shape[:2]
        d = np.zeros((H, W), np.float32)
        _check(self.lib.hbp_plane_to_disp(self.h, np.ascontiguousarray(plane3, np.float32), W, H, d))
        return d


class Volume(_Handle):
    _free_fn = "hbp_volume_free"

    def __init__(self, ctx: Ctx, left, right, D, right_view=False, f32=None, cost="Census"):
        self.ctx = ctx
        h = C.c_void_p()
        if f32 is not None:
            D, H, W = f32.shape
            _check(ctx.lib.hbp_cost_volume_f32(ctx.h, np.ascontiguousarray(f32, np.float32), W, H, D, b"MC-CNN",
                                               C.byref(h)))
        else:
            H, W = left.shape[:2]
            _check(ctx.lib.hbp_cost_volume(ctx.h, np.ascontiguousarray(left), np.ascontiguousarray(right), W, H, D,
                                           cost.encode(), C.byref(h)))
        self.f32 = f32 is not None or cost == "AD"
        self._adopt(ctx, h)
        self.W, self.H, self.D = W, H, D
        if right_view:
            _check(ctx.lib.hbp_cost_to_right(ctx.h, self.h))

    def download_f32(self):
        vol = np.zeros((self.D, self.H, self.W), np.float32)
        wta = np.zeros((self.H, self.W), np.float32)
        _check(self.ctx.lib.hbp_volume_download_f32(self.ctx.h, self.h, _np_ptr(vol), _np_ptr(wta)))
        return vol, wta

    def download(self):
        vol = np.zeros((self.D, self.H, self.W), np.uint8)
        wta = np.zeros((self.H, self.W), np.float32)
        _check(self.ctx.lib.hbp_volume_download(self.ctx.h, self.h, _np_ptr(vol), _np_ptr(wta)))
        return vol, wta


class Pyramid(_Handle):
    _free_fn = "hbp_pyramid_free"

    def __init__(self, ctx: Ctx, bgr, num_nb=4, link=2, target_clus=1, sigma=0.0):
        self.ctx = ctx
        H, W = bgr.shape[:2]
        h = C.c_void_p()
        if target_clus == 1:
            _check(ctx.lib.hbp_pyramid_build(ctx.h, np.ascontiguousarray(bgr), W, H, link, num_nb, float(sigma),
                                             C.byref(h)))
        else:
            _check(ctx.lib.hbp_pyramid_build_target(ctx.h, np.ascontiguousarray(bgr), W, H, link, num_nb,
                                                    float(sigma), target_clus, C.byref(h)))
        self._adopt(ctx, h)
        self.W, self.H = W, H

    @property
    def levels(self):
        return self.ctx.lib.hbp_pyramid_levels(self.h)

    def level_size(self, l):
        """(nodes, directed edges) of level l"""
        V, E = C.c_int(), C.c_int()
        _check(self.ctx.lib.hbp_pyramid_level_size(self.h, l, C.byref(V), C.byref(E)))
        return V.value, E.value

    def level(self, l):
        V, E = C.c_int(), C.c_int()
        _check(self.ctx.lib.hbp_pyramid_level_size(self.h, l, C.byref(V), C.byref(E)))
        V, E = V.value, E.value
        clus = np.zeros(V, np.int32)
        mean = np.zeros(6 * V, np.float32)
        bd = np.zeros(2 * E, np.int32)
        bdl = np.zeros(E, np.int32)
        _check(self.ctx.lib.hbp_pyramid_download_level(self.ctx.h, self.h, l, _np_ptr(clus), _np_ptr(mean),
                                                       _np_ptr(bd), _np_ptr(bdl)))
        return dict(V=V, E=E, clus=clus, mean=mean.reshape(6, V), bd=bd.reshape(E, 2), bdl=bdl)


class BP(_Handle):
    _free_fn = "hbp_bp_free"

    def __init__(self, ctx: Ctx, pyr: Pyramid, vol: Volume, alpha=1000.1, tau_s=0.6, min_level=2, use_cost_vol=1):
        self.ctx, self.pyr, self.vol = ctx, pyr, vol
        h = C.c_void_p()
        _check(ctx.lib.hbp_bp_create(ctx.h, pyr.h, vol.h, alpha, tau_s, min_level, use_cost_vol, C.byref(h)))
        self._adopt(ctx, h)
        self.W, self.H = pyr.W, pyr.H

    @property
    def total_sweeps(self):
        return self.ctx.lib.hbp_bp_total_sweeps(self.h)

    @property
    def level(self):
        return self.ctx.lib.hbp_bp_level(self.h)

    def step(self, n=1 << 30):
        done = C.c_int()
        _check(self.ctx.lib.hbp_bp_step(self.ctx.h, self.h, n, C.byref(done)))
        return done.value

    def run(self):
        disp = np.zeros((self.H, self.W), np.float32)
        plane = np.zeros((self.H, self.W, 3), np.float32)
        _check(self.ctx.lib.hbp_bp_run(self.ctx.h, self.h, _np_ptr(disp), _np_ptr(plane)))
        return disp, plane

    def state(self):
        l = self.level
        V, E = C.c_int(), C.c_int()
        _check(self.ctx.lib.hbp_bp_level_size(self.h, l, C.byref(V), C.byref(E)))
        V, E = V.value, E.value
        planes = np.zeros(12 * V * 4, np.float32)
        cost = np.zeros(12 * V, np.float32)
        belief = np.zeros(12 * V, np.float32)
        msg = np.zeros(E * 8, np.float32)
        _check(self.ctx.lib.hbp_bp_get_state(self.ctx.h, self.h, _np_ptr(planes), _np_ptr(cost), _np_ptr(belief),
                                             _np_ptr(msg)))
        return dict(level=l, V=V, E=E, planes=planes.reshape(12, V, 4), cost=cost.reshape(12, V),
                    belief=belief.reshape(12, V), msg=msg.reshape(E, 8))


def exported_symbols():
    """Names declared in include/hbp_hip.h (for the export check)."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(hbp_[a-z0-9_]+)\s*\(", txt)))


class Server(_Handle):
    """Continuous-batching request server on a context (hbp_server_*, hbp_submit, hbp_poll)."""

    def __init__(self, ctx: Ctx, max_active=64, hbm_budget=64 << 30, alpha=1000.1, tau_s=0.6):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.lib.hbp_server_create(ctx.h, max_active, int(hbm_budget), alpha, tau_s, C.byref(h)))
        self._adopt(ctx, h)
        self.sizes = {}

    def submit(self, left, right, labels, min_level=2):
        left, right = np.ascontiguousarray(left), np.ascontiguousarray(right)
        H, W = left.shape[:2]
        t = C.c_longlong()
        _check(self.ctx.lib.hbp_submit(self.h, left, right, W, H, labels, min_level, C.byref(t)))
        self.sizes[t.value] = (H, W)
        return t.value

    def step(self):
        """(pairs on the device, queued, completed by this step)"""
        a, p, d = C.c_int(), C.c_int(), C.c_int()
        _check(self.ctx.lib.hbp_server_step(self.h, C.byref(a), C.byref(p), C.byref(d)))
        return a.value, p.value, d.value

    def poll(self, ticket):
        """(disp, plane) when done, else None; raises for a failed or cancelled request."""
        H, W = self.sizes.get(ticket, (0, 0))
        disp = np.empty((H, W), np.float32)          # written only once the request is done
        plane = np.empty((H, W, 3), np.float32)
        r = self.ctx.lib.hbp_poll(self.h, ticket, _np_ptr(disp), _np_ptr(plane))
        if r == 1:
            return None
        if r != 0:
            raise HbpError(self.ctx.lib.hbp_last_error().decode())
        del self.sizes[ticket]
        return disp, plane

    def set_max_pending(self, n):
        _check(self.ctx.lib.hbp_server_set_max_pending(self.h, int(n)))

    def stats(self):
        """(device bytes accounted now, at most, submissions refused)"""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        _check(self.ctx.lib.hbp_server_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def cancel(self, ticket):
        _check(self.ctx.lib.hbp_cancel(self.h, ticket))
        self.sizes.pop(ticket, None)

    def free(self):
        h = getattr(self, "h", None)
        if not h:
            return
        self.h = None
        self.ctx._children.discard(self)
        if self.ctx.h and _alive():
            self.ctx.lib.hbp_server_destroy(h)

    close = free
