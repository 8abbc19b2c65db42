"""The drop-in classes' surfaces that no other GPU test reaches, through bin/api_check (built
from tests/api_check.cpp): SegHAC::set_frame and set_seq_img (segment.hpp:87-108) followed by
run_ms, each pyramid equal to the oracle's for the new image, and StereoCost's float getter
g_img_cost_d (stereo_cost.hpp:42-46) for the Census cost, equal to the oracle's volume; plus the
same f32 volume (both views) through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hbp-isp_amd", "bin", "api_check")


def _levels(path_fmt, po):
    for l in range(po.levels):
        lv = po.level(l)
        raw = np.fromfile(path_fmt.format(l), np.int32)
        assert raw.size == lv["V"] + 2 * lv["E"], l
        assert np.array_equal(raw[:lv["V"]], lv["clus"]), l
        assert np.array_equal(raw[lv["V"]:].reshape(-1, 2), lv["bd"]), l
    assert not os.path.exists(path_fmt.format(po.levels))


def test_video_hooks_and_float_volume_getter(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("bin/api_check is not built (make -C hbp-isp_amd)")
    W, H, D = 120, 90, 24
    a, b, _ = synth.make_pair(51, W, H, D, (3, 8))
    t, _, _ = synth.make_pair(52, H, W, D, (3, 8))     # W rows, H columns: the transposed frame
    for name, img in (("a", a), ("b", b), ("t", t)):
        np.ascontiguousarray(img).tofile(str(tmp_path / f"{name}.bgr"))
    print("[case] api_check 120x90 D24", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path), str(W), str(H), str(D), str(tmp_path / "a.bgr"),
                        str(tmp_path / "b.bgr"), str(tmp_path / "t.bgr")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _levels(str(tmp_path / "seg_a_{}.bin"), O.Pyramid(a))
    _levels(str(tmp_path / "seg_b_{}.bin"), O.Pyramid(b))   # set_frame
    _levels(str(tmp_path / "seg_t_{}.bin"), O.Pyramid(t))   # set_seq_img, transposed
    vo, wo = O.cost_volume(a, b, D)
    vol = np.fromfile(str(tmp_path / "vol.bin"), np.float32).reshape(D, H, W)
    wta = np.fromfile(str(tmp_path / "wta.bin"), np.float32).reshape(H, W)
    assert np.array_equal(vol, vo.astype(np.float32)) and np.array_equal(wta, wo)


@pytest.mark.parametrize("right_view", [False, True])
def test_census_volume_as_f32_through_the_c_abi(ctx, right_view):
    W, H, D = 160, 120, 48
    left, right, _ = synth.make_pair(53, W, H, D, (3, 8))
    vo, wo = O.cost_volume(left, right, D, right_view=right_view)
    vol = ctx.cost_volume(left, right, D, right_view=right_view)
    vf, wf = vol.download_f32()
    assert np.array_equal(vf, vo.astype(np.float32)) and np.array_equal(wf, wo)
    vol.free()
