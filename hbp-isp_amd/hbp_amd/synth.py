"""Seeded synthetic stereo pairs: layered slanted planes with value-noise texture.

Scene model (deterministic from ``seed``):
  * ``n_planes`` slanted planes ``d = a*(x-cx) + b*(y-cy) + c`` with ``|a|,|b| < 0.05`` and
    ``c ~ U[4, 60] * D/64`` (so the 450x375/D=64 case draws c from [4, 60]).  Plane 0 is the
    background and covers the whole image; every other plane owns an ellipse or a rectangle.
  * Visibility is a z-buffer on disparity: the nearest (largest d) plane that covers a point
    wins, in the left view at (x, y) and in the right view at the left-image point that maps
    to (x_r, y), i.e. ``x_l = x_r + d(x_l, y)``.  Occlusions follow from the layering.
  * Texture: per plane and channel, 3-octave value noise (lattice pitch 24/12/6 px, weights
    1, 1/2, 1/4, smoothstep interpolation) around a random base colour, sampled at the left-
    image coordinates of the surface point, plus per-pixel uniform noise in [-2, 2] grey
    levels, drawn independently for the two views.
  * The right view gets gain ~ U[0.95, 1.05] and bias ~ U[-5, 5]; both views are rounded and
    clamped to uint8.  Images are BGR, interleaved, row-major (the cv::Mat CV_8UC3 layout).

Random numbers come from SplitMix64 used as a counter-based hash (Steele, Lea, Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): value = mix(seed_key +
GOLDEN * (counter + 1)), so every draw is addressed by (seed, stream, counter) and the
generator is vectorised with numpy without any hidden state.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def _key(seed: int, *stream: int) -> np.uint64:
    k = _mix(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
    for s in stream:
        with np.errstate(over="ignore"):
            k = _mix(k ^ (np.uint64(s & 0xFFFFFFFFFFFFFFFF) + _GOLDEN))
    return np.uint64(k)


def _u01(key: np.uint64, counter) -> np.ndarray:
    """Uniform [0, 1) doubles addressed by an integer counter array."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix(key + _GOLDEN * (c + np.uint64(1)))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _value_noise(key: np.uint64, x: np.ndarray, y: np.ndarray, pitch: float) -> np.ndarray:
    """Smooth value noise in [-1, 1] with lattice pitch ``pitch`` px."""
    fx = x / pitch
    fy = y / pitch
    gx = np.floor(fx)
    gy = np.floor(fy)
    tx = fx - gx
    ty = fy - gy
    tx = tx * tx * (3.0 - 2.0 * tx)
    ty = ty * ty * (3.0 - 2.0 * ty)
    gx = gx.astype(np.int64) + 4096
    gy = gy.astype(np.int64) + 4096
    if gx.size == 0:
        return np.zeros(gx.shape)
    # the lattice values once per lattice point of the covering box, then gathered per pixel
    x0, y0 = int(gx.min()), int(gy.min())
    ix = np.arange(x0, int(gx.max()) + 2, dtype=np.uint64)
    iy = np.arange(y0, int(gy.max()) + 2, dtype=np.uint64)
    lat = _u01(key, (iy[:, None] << np.uint64(20)) | ix[None, :]) * 2.0 - 1.0
    gx -= x0
    gy -= y0
    v00 = lat[gy, gx]
    v10 = lat[gy, gx + 1]
    v01 = lat[gy + 1, gx]
    v11 = lat[gy + 1, gx + 1]
    top = v00 + (v10 - v00) * tx
    bot = v01 + (v11 - v01) * tx
    return top + (bot - top) * ty


class Scene:
    """Plane parameters of one synthetic scene (left-image coordinates)."""

    def __init__(self, seed: int, width: int, height: int, ndisp: int, n_planes: tuple[int, int]):
        self.seed, self.width, self.height, self.ndisp = seed, width, height, ndisp
        k = _key(seed, 1)
        lo, hi = n_planes
        n = lo + int(_u01(k, 0) * (hi - lo + 1))
        n = min(max(n, lo), hi)
        cnt = np.arange(1, 1 + 16 * n, dtype=np.uint64)
        r = _u01(k, cnt).reshape(n, 16)
        scale = ndisp / 64.0
        self.a = (r[:, 0] * 2.0 - 1.0) * 0.05
        self.b = (r[:, 1] * 2.0 - 1.0) * 0.05
        self.c = (4.0 + r[:, 2] * 56.0) * scale
        self.cx = r[:, 3] * width
        self.cy = r[:, 4] * height
        self.rx = (0.08 + 0.30 * r[:, 5]) * width
        self.ry = (0.08 + 0.30 * r[:, 6]) * height
        self.ellipse = r[:, 7] < 0.5
        self.base = 40.0 + 175.0 * r[:, 8:11]          # B, G, R base colour
        self.amp = 25.0 + 45.0 * r[:, 11:14]
        self.n = n
        # background: far, whole image
        self.c[0] = 4.0 * scale * (0.5 + 0.5 * r[0, 2])
        self.rx[0] = self.ry[0] = -1.0

    def disparity(self, k: int, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        return self.a[k] * (x - self.cx[k]) + self.b[k] * (y - self.cy[k]) + self.c[k]

    def covers(self, k: int, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        if self.rx[k] < 0:
            return np.ones(np.broadcast(x, y).shape, dtype=bool)
        dx = (x - self.cx[k]) / self.rx[k]
        dy = (y - self.cy[k]) / self.ry[k]
        if self.ellipse[k]:
            return dx * dx + dy * dy <= 1.0
        return (np.abs(dx) <= 1.0) & (np.abs(dy) <= 1.0)

    def texture(self, k: int, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        out = np.empty(x.shape + (3,), dtype=np.float64)
        for ch in range(3):
            acc = np.zeros(x.shape, dtype=np.float64)
            for o, (pitch, w) in enumerate(((24.0, 1.0), (12.0, 0.5), (6.0, 0.25))):
                acc += w * _value_noise(_key(self.seed, 2, k, ch, o), x, y, pitch)
            out[..., ch] = self.base[k, ch] + self.amp[k, ch] * acc / 1.75
        return out


def make_pair(seed: int, width: int, height: int, ndisp: int,
              n_planes: tuple[int, int] = (6, 12), rows: tuple[int, int] | None = None):
    """Return (left_bgr u8 [H,W,3], right_bgr u8 [H,W,3], gt_disp f32 [H,W]); with rows = (y0, y1)
    only those rows of the same pair (every pixel depends on its own coordinates alone)."""
    sc = Scene(seed, width, height, ndisp, n_planes)
    y0, y1 = (0, height) if rows is None else rows
    ys, xs = np.mgrid[y0:y1, 0:width].astype(np.float64)
    hh = y1 - y0

    # left view: z-buffer on disparity
    best_d = np.full((hh, width), -np.inf)
    best_k = np.zeros((hh, width), dtype=np.int32)
    for k in range(sc.n):
        d = sc.disparity(k, xs, ys)
        m = sc.covers(k, xs, ys) & (d > best_d)
        best_d[m] = d[m]
        best_k[m] = k
    left = np.empty((hh, width, 3), dtype=np.float64)
    for k in range(sc.n):
        m = best_k == k
        if m.any():
            left[m] = sc.texture(k, xs[m], ys[m])

    # right view: x_l = (x_r + a*cx - b*(y-cy) ... ) solved per plane
    rbest_d = np.full((hh, width), -np.inf)
    rbest_k = np.zeros((hh, width), dtype=np.int32)
    rbest_xl = np.zeros((hh, width))
    for k in range(sc.n):
        xl = (xs - sc.a[k] * sc.cx[k] + sc.b[k] * (ys - sc.cy[k]) + sc.c[k]) / (1.0 - sc.a[k])
        d = sc.disparity(k, xl, ys)
        m = sc.covers(k, xl, ys) & (d > rbest_d)
        rbest_d[m] = d[m]
        rbest_k[m] = k
        rbest_xl[m] = xl[m]
    right = np.empty((hh, width, 3), dtype=np.float64)
    for k in range(sc.n):
        m = rbest_k == k
        if m.any():
            right[m] = sc.texture(k, rbest_xl[m], ys[m])

    kn = _key(seed, 3)
    npx = height * width * 3
    a0, a1 = y0 * width * 3, y1 * width * 3
    left += _u01(kn, np.arange(a0, a1, dtype=np.uint64)).reshape(hh, width, 3) * 4.0 - 2.0
    right += _u01(kn, np.arange(npx + a0, npx + a1, dtype=np.uint64)).reshape(hh, width, 3) * 4.0 - 2.0
    kg = _key(seed, 4)
    gain = 0.95 + 0.10 * float(_u01(kg, 0))
    bias = -5.0 + 10.0 * float(_u01(kg, 1))
    right = right * gain + bias

    left_u8 = np.clip(np.rint(left), 0, 255).astype(np.uint8)
    right_u8 = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    gt = np.clip(best_d, 0.0, ndisp - 1).astype(np.float32)
    return left_u8, right_u8, gt


def _rows_job(a):
    return make_pair(*a)


def make_pair_parallel(seed: int, width: int, height: int, ndisp: int,
                       n_planes: tuple[int, int] = (6, 12), workers: int = 16, pool=None):
    """make_pair's result, its rows generated in chunks by a process pool (`pool`, or one of
    `workers` processes made for the call)."""
    import multiprocessing as mp
    n = max(1, height // 64)
    cuts = [height * i // n for i in range(n + 1)]
    jobs = [(seed, width, height, ndisp, n_planes, (cuts[i], cuts[i + 1])) for i in range(n)]
    if pool is not None:
        parts = pool.map(_rows_job, jobs)
    else:
        with mp.get_context("fork").Pool(max(1, min(workers, n))) as own:
            parts = own.map(_rows_job, jobs)
    return tuple(np.concatenate([p[i] for p in parts], axis=0) for i in range(3))


def make_pairs_parallel(specs, n_planes: tuple[int, int] = (6, 12), workers: int = 16, pool=None):
    """Pairs (seed, width, height, ndisp) generated by a process pool: [(left, right), ...]."""
    import multiprocessing as mp
    jobs = []
    for k, (sd, w, h, d) in enumerate(specs):
        n = max(1, h // 64)
        jobs += [(k, (sd, w, h, d, n_planes, (h * i // n, h * (i + 1) // n))) for i in range(n)]
    if pool is not None:
        parts = pool.map(_rows_job, [j for _, j in jobs])
    else:
        with mp.get_context("fork").Pool(max(1, min(workers, len(jobs)))) as own:
            parts = own.map(_rows_job, [j for _, j in jobs])
    out = []
    for k in range(len(specs)):
        mine = [p for (kk, _), p in zip(jobs, parts) if kk == k]
        out.append(tuple(np.concatenate([p[i] for p in mine], axis=0) for i in range(2)))
    return out


# Workloads named by what they are (width, height, max disparity, plane count range).
PLAIN_SMALL = dict(width=450, height=375, ndisp=64, n_planes=(6, 12))
PLAIN_LARGE = dict(width=1472, height=984, ndisp=256, n_planes=(20, 40))


def mixed_batch(seed: int, count: int = 64):
    """Per-pair shapes of the uneven batch: W~U[320,1920], H~U[240,1080] (multiples of 8),
    D drawn from {64,128,192,256}; returns a list of (seed_i, W, H, D)."""
    k = _key(seed, 5)
    r = _u01(k, np.arange(3 * count, dtype=np.uint64)).reshape(count, 3)
    out = []
    for i in range(count):
        w = int(320 + r[i, 0] * (1920 - 320)) // 8 * 8
        h = int(240 + r[i, 1] * (1080 - 240)) // 8 * 8
        d = (64, 128, 192, 256)[min(int(r[i, 2] * 4), 3)]
        out.append((seed * 1000 + i, max(w, 320), max(h, 240), d))
    return out
