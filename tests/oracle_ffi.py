"""ctypes bindings to the oracle (oracle/lib/libhbp_oracle.so): test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "lib", "libhbp_oracle.so")

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_pin_cbrt": (C.c_double, [C.c_double]),
        "orc_pin_exp": (C.c_double, [C.c_double]),
        "orc_pin_pow24": (C.c_double, [C.c_double]),
        "orc_gauss_weights": (C.c_int, [C.c_double, _f32p]),
        "orc_gauss_blur": (C.c_int, [_u8p, C.c_int, C.c_int, C.c_double, _f32p]),
        "orc_lab_f": (None, [_f32p, C.c_long, _f32p]),
        "orc_pyramid_build_sigma": (C.c_void_p, [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
        "orc_i64_to_f32": (C.c_float, [C.c_int64]),
        "orc_f2i_rd": (C.c_int, [C.c_float]),
        "orc_fixed_shift": (C.c_int, [C.c_float, C.c_float, C.c_int, C.c_int]),
        "orc_amin_fold": (C.c_float, [C.c_float, C.c_float]),
        "orc_grey": (None, [_u8p, C.c_int, C.c_int, _u8p]),
        "orc_census": (None, [_u8p, C.c_int, C.c_int, _u32p]),
        "orc_hamming": (None, [_u32p, _u32p, C.c_int, C.c_int, C.c_int, C.c_int, _u8p]),
        "orc_wta": (None, [_u8p, C.c_int, C.c_int, C.c_int, _f32p]),
        "orc_vol_to_right": (None, [_u8p, C.c_int, C.c_int, C.c_int]),
        "orc_cost_volume": (C.c_int, [_u8p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, _u8p, _f32p]),
        "orc_lab": (None, [_u8p, C.c_int, C.c_int, _f32p]),
        "orc_nn_grid": (None, [_f32p, C.c_int, C.c_int, C.c_int, _i32p, _f32p, _i32p]),
        "orc_cc_max": (None, [_i32p, C.c_int, _i32p]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_crtrees_literal": (None, [_i32p, C.c_int, C.c_int, _i32p]),
        "orc_compact": (C.c_int, [_i32p, C.c_int]),
        "orc_cost_volume_ad": (C.c_int, [_u8p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, _f32p, _f32p]),
        "orc_pyramid_build": (C.c_void_p, [_u8p, C.c_int, C.c_int, C.c_int]),
        "orc_pyramid_build_link": (C.c_void_p, [_u8p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_pyramid_build_target": (C.c_void_p, [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_pyr_levels": (C.c_int, [C.c_void_p]),
        "orc_pyr_level_size": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_pyr_get_level": (None, [C.c_void_p, C.c_int, _i32p, _f32p, _i32p, _i32p]),
        "orc_pyr_free": (None, [C.c_void_p]),
        "orc_bp_new": (C.c_void_p, [C.c_void_p, _u8p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                    C.c_float, C.c_float, C.c_int, C.c_int]),
        "orc_bp_new_f32": (C.c_void_p, [C.c_void_p, _f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                        C.c_float, C.c_float, C.c_int, C.c_int]),
        "orc_cost_volume_f32": (C.c_int, [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, _f32p, _f32p]),
        "orc_bp_step": (C.c_int, [C.c_void_p, C.c_int]),
        "orc_bp_level": (C.c_int, [C.c_void_p]),
        "orc_bp_sweeps_done": (C.c_int, [C.c_void_p]),
        "orc_bp_level_size": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_bp_get_state": (None, [C.c_void_p, _f32p, _f32p, _f32p, _f32p]),
        "orc_bp_map": (None, [C.c_void_p, _f32p, _f32p]),
        "orc_bp_free": (None, [C.c_void_p]),
        "orc_bp_total_sweeps": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "orc_lr_consistency": (None, [_f32p, _f32p, C.c_int, C.c_int, _u8p, _u8p]),
        "orc_dilate3": (None, [_u8p, C.c_int, C.c_int, _u8p]),
        "orc_lr_fill": (None, [_u8p, _u8p, _f32p, C.c_int, C.c_int]),
        "orc_plane_to_disp": (None, [_f32p, C.c_int, C.c_int, _f32p]),
        "orc_left_right_check": (None, [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, _u8p, _u8p, _u8p]),
        "orc_wm_weight": (C.c_float, [C.c_float]),
        "orc_weighted_median": (None, [_u8p, _u8p, _f32p, C.c_int, C.c_int]),
        "orc_median7": (None, [_f32p, C.c_int, C.c_int]),
        "orc_eval_bad_pixels": (None, [_f32p, _f32p, _u8p, C.c_int, C.c_int, C.c_double, _i64p]),
        "orc_plain_pair": (C.c_int, [_u8p, _u8p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_int, C.c_int, _f32p, _f32p, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def cost_volume(left, right, D, right_view=False):
    H, W = left.shape[:2]
    vol = np.zeros((D, H, W), np.uint8)
    dwta = np.zeros((H, W), np.float32)
    rc = lib.orc_cost_volume(np.ascontiguousarray(left), np.ascontiguousarray(right), W, H, D,
                             int(right_view), vol, dwta)
    assert rc == 0
    return vol, dwta


class Pyramid:
    def __init__(self, bgr, num_nb=4, link=2, target_clus=1, sigma=0.0):
        H, W = bgr.shape[:2]
        self.W, self.H = W, H
        self.h = lib.orc_pyramid_build_sigma(np.ascontiguousarray(bgr), W, H, num_nb, link, target_clus, sigma)
        assert self.h, "bad linkage, target_clus or sigma"

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_pyr_free(self.h)
            self.h = None

    @property
    def levels(self):
        return lib.orc_pyr_levels(self.h)

    def level(self, l):
        V, E = C.c_int(), C.c_int()
        lib.orc_pyr_level_size(self.h, l, C.byref(V), C.byref(E))
        V, E = V.value, E.value
        clus = np.zeros(V, np.int32)
        mean = np.zeros(6 * V, np.float32)
        bd = np.zeros(2 * E, np.int32)
        bdl = np.zeros(E, np.int32)
        lib.orc_pyr_get_level(self.h, l, clus, mean, bd, bdl)
        return dict(V=V, E=E, clus=clus, mean=mean.reshape(6, V), bd=bd.reshape(E, 2), bdl=bdl)


def cost_volume_ad(left, right, D, right_view=False):
    H, W = left.shape[:2]
    vol = np.zeros((D, H, W), np.float32)
    dwta = np.zeros((H, W), np.float32)
    assert lib.orc_cost_volume_ad(np.ascontiguousarray(left), np.ascontiguousarray(right), W, H, D,
                                  int(right_view), vol, dwta) == 0
    return vol, dwta


def cost_volume_f32(raw, tau=0.5, right_view=False):
    D, H, W = raw.shape
    vol = np.zeros((D, H, W), np.float32)
    dwta = np.zeros((H, W), np.float32)
    assert lib.orc_cost_volume_f32(np.ascontiguousarray(raw, np.float32), W, H, D, tau, int(right_view), vol,
                                   dwta) == 0
    return vol, dwta


class BP:
    def __init__(self, pyr: Pyramid, vol, dwta, D, alpha=1000.1, tau_s=0.6, min_level=2,
                 use_cost_vol=1, tau=None):
        self.pyr, self.vol, self.dwta = pyr, np.ascontiguousarray(vol), np.ascontiguousarray(dwta)
        self.W, self.H, self.D = pyr.W, pyr.H, D
        if self.vol.dtype == np.float32:
            self.h = lib.orc_bp_new_f32(pyr.h, self.vol, self.dwta, pyr.W, pyr.H, D, 0.5 if tau is None else tau,
                                        alpha, tau_s, min_level, use_cost_vol)
        else:
            self.h = lib.orc_bp_new(pyr.h, self.vol, self.dwta, pyr.W, pyr.H, D, 25.0 if tau is None else tau,
                                    alpha, tau_s, min_level, use_cost_vol)
        assert self.h, "invalid min_level"

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_bp_free(self.h)
            self.h = None

    def step(self, n=1 << 30):
        return lib.orc_bp_step(self.h, n)

    @property
    def level(self):
        return lib.orc_bp_level(self.h)

    def state(self):
        l = self.level
        V, E = C.c_int(), C.c_int()
        lib.orc_bp_level_size(self.h, l, C.byref(V), C.byref(E))
        V, E = V.value, E.value
        planes = np.zeros(12 * V * 4, np.float32)
        cost = np.zeros(12 * V, np.float32)
        belief = np.zeros(12 * V, np.float32)
        msg = np.zeros(E * 8, np.float32)
        lib.orc_bp_get_state(self.h, planes, cost, belief, msg)
        return dict(level=l, V=V, E=E, planes=planes.reshape(12, V, 4), cost=cost.reshape(12, V),
                    belief=belief.reshape(12, V), msg=msg.reshape(E, 8))

    def map(self):
        disp = np.zeros((self.H, self.W), np.float32)
        plane3 = np.zeros((self.H, self.W, 3), np.float32)
        lib.orc_bp_map(self.h, disp, plane3)
        return disp, plane3


def plain_pair(left, right, D, alpha=1000.1, tau_s=0.6, min_level=2, bp_levels=0):
    H, W = left.shape[:2]
    disp = np.zeros((H, W), np.float32)
    plane3 = np.zeros((H, W, 3), np.float32)
    L = C.c_int()
    ml = lib.orc_plain_pair(np.ascontiguousarray(left), np.ascontiguousarray(right), W, H, D, alpha,
                            tau_s, min_level, bp_levels, disp, plane3, C.byref(L))
    return disp, plane3, L.value, ml


def eval_bad_pixels(disp, gt, nonocc, threshold):
    """[(valid, good), (nonocc, good), (occ, good)] as the oracle counts them."""
    H, W = gt.shape
    c = np.zeros(6, np.int64)
    lib.orc_eval_bad_pixels(np.ascontiguousarray(disp, np.float32), np.ascontiguousarray(gt, np.float32),
                            np.ascontiguousarray(nonocc, np.uint8), W, H, float(threshold), c)
    return [(int(c[0]), int(c[1])), (int(c[2]), int(c[3])), (int(c[4]), int(c[5]))]


def gauss_weights(sigma):
    w = np.zeros(289, np.float32)
    S = lib.orc_gauss_weights(sigma, w)
    assert S >= 0, "round(3 sigma) > 8"
    F = 2 * S + 1
    return S, w[:F * F].reshape(F, F)


def gauss_blur(bgr, sigma):
    H, W = bgr.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    S = lib.orc_gauss_blur(np.ascontiguousarray(bgr), W, H, sigma, out)
    return S, out


def lab_f(img4):
    img4 = np.ascontiguousarray(img4, np.float32)
    out = np.zeros_like(img4)
    lib.orc_lab_f(img4, img4.size // 4, out)
    return out
