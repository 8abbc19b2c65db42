"""Middlebury scene files and the bad-pixel evaluation, for Python callers of the library.

Formats (as HBP-ISP's benchmark driver reads and writes them, main_bench_middv3.cu:12-48):
  * PFM (cvutils.hpp:23-144): "Pf" one channel / "PF" three, "<w> <h>", a scale whose sign gives
    the byte order (negative = little endian), then float32 rows from the bottom row up.
  * calib.txt (ArgsParser.h:49-71): "name = value" lines, cam0/cam1 as 3x3 matrices.
  * PNG: 8-bit non-interlaced images (grey, grey+alpha, RGB, RGBA, palette) via zlib; images
    are returned in BGR order like cv::imread.
The evaluation (Evaluator.h:46-101) counts on the device through the C ABI
(hbp_eval_bad_pixels): bad-pixel rates in percent over all valid pixels (GT > 0 and GT != +inf),
the non-occluded mask and the occluded pixels (valid and not in the mask), a pixel being bad
when |disp - GT| > threshold in float.
"""
import os
import re
import struct
import zlib

import numpy as np

__all__ = ["read_pfm", "write_pfm", "read_calib", "read_png", "write_png", "error_threshold", "evaluate",
           "load_scene"]


def read_pfm(path):
    """float32 [H][W] or [H][W][3], top row first; None when the file is absent."""
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        data = f.read()
    tokens = []
    pos = 0
    while len(tokens) < 4:
        m = re.compile(rb"\s*(\S+)").match(data, pos)
        if not m:
            raise ValueError(f"{path}: truncated PFM header")
        tokens.append(m.group(1))
        pos = m.end()
    magic, w, h, scale = tokens[0], int(tokens[1]), int(tokens[2]), float(tokens[3])
    ch = {b"Pf": 1, b"PF": 3}.get(magic)
    if ch is None or w <= 0 or h <= 0:
        raise ValueError(f"{path}: not a PFM file")
    n = w * h * ch
    if len(data) - (pos + 1) < 4 * n:
        raise ValueError(f"{path}: truncated PFM data")
    # the data are the file's last w*h*ch floats (the header ends with one whitespace byte)
    arr = np.frombuffer(data[len(data) - 4 * n:], dtype="<f4" if scale < 0 else ">f4").astype(np.float32)
    arr = arr.reshape((h, w, ch) if ch == 3 else (h, w))
    return np.ascontiguousarray(arr[::-1])


def write_pfm(path, img):
    """Writes float32 [H][W] ("Pf") or [H][W][3] ("PF") the way save_pfm_file does: little endian,
    scale -1/255, rows from the bottom up."""
    img = np.asarray(img, np.float32)
    ch = 1 if img.ndim == 2 else img.shape[2]
    if ch not in (1, 3):
        raise ValueError("write_pfm: 1 or 3 channels")
    with open(path, "wb") as f:
        f.write(b"%s\n%d %d\n%f\n" % (b"Pf" if ch == 1 else b"PF", img.shape[1], img.shape[0], -1.0 / 255.0))
        f.write(np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())


# calib.txt is read line by line in a fixed order, each line with its own scanf format
# (ArgsParser.h:49-71); a line that does not match leaves its fields at 0.
_CALIB_LINES = [("cam0", "cam0 = [%f %f %f; %f %f %f; %f %f %f]"), ("cam1", "cam1 = [%f %f %f; %f %f %f; %f %f %f]"),
                ("doffs", "doffs = %f"), ("baseline", "baseline = %f"), ("width", "width = %d"),
                ("height", "height = %d"), ("ndisp", "ndisp = %d"), ("isint", "isint = %d"), ("vmin", "vmin = %d"),
                ("vmax", "vmax = %d"), ("dyavg", "dyavg = %f"), ("dymax", "dymax = %f")]
_FLOAT = re.compile(r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.I)
_INT = re.compile(r"[+-]?\d+")


def _scan(fmt, text):
    """The values scanf(fmt) assigns from text before its first mismatch (%f, %d, whitespace
    and literal characters only)."""
    out, i, k = [], 0, 0
    while k < len(fmt):
        if fmt[k].isspace():
            while i < len(text) and text[i].isspace():
                i += 1
            k += 1
        elif fmt[k] == "%":
            while i < len(text) and text[i].isspace():
                i += 1
            m = (_FLOAT if fmt[k + 1] == "f" else _INT).match(text, i)
            if not m:
                break
            out.append(np.float32(m.group(0)) if fmt[k + 1] == "f" else int(m.group(0)))
            i, k = m.end(), k + 2
        elif i < len(text) and text[i] == fmt[k]:
            i, k = i + 1, k + 1
        else:
            break
    return out


def read_calib(path):
    """dict of calib.txt's fields (cam0, cam1 as float32 [3][3], floats as float32, ints as
    int); None when the file is absent."""
    if not os.path.exists(path):
        return None
    out = {name: (np.zeros((3, 3), np.float32) if name.startswith("cam") else
                  (np.float32(0) if "%f" in fmt else 0)) for name, fmt in _CALIB_LINES}
    with open(path) as f:
        lines = f.readlines()
    for (name, fmt), line in zip(_CALIB_LINES, lines):
        vals = _scan(fmt, line)
        if name.startswith("cam"):
            out[name].reshape(-1)[:len(vals)] = vals
        elif vals:
            out[name] = vals[0]
    return out


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def read_png(path, color=True):
    """8-bit non-interlaced PNG -> u8 [H][W][3] BGR (color) or [H][W] (grey); None when absent."""
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, palette, hdr = 8, [], None, None
    while pos + 12 <= len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body[:13])
        elif kind == b"PLTE":
            palette = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
        pos += 12 + n
    W, H, depth, ctype, _, _, interlace = hdr
    nch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}.get(ctype)
    if depth != 8 or nch is None or interlace:
        raise ValueError(f"{path}: only 8-bit non-interlaced PNG images are read")
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8).reshape(H, W * nch + 1)
    img = np.zeros((H, W * nch), np.uint8)
    prev = np.zeros(W * nch, np.int32)
    for y in range(H):
        ft, src = raw[y, 0], raw[y, 1:].astype(np.int32)
        if ft == 0:
            row = src
        elif ft == 1:     # Sub: a running sum per channel
            row = np.cumsum(src.reshape(W, nch), axis=0).reshape(-1) & 255
        elif ft == 2:     # Up
            row = (src + prev) & 255
        elif ft in (3, 4):  # Average, Paeth: serial along the row
            r, up, sv = [0] * (W * nch), prev.tolist(), src.tolist()
            for i in range(W * nch):
                a = r[i - nch] if i >= nch else 0
                b = up[i]
                if ft == 3:
                    pred = (a + b) >> 1
                else:
                    c = up[i - nch] if i >= nch else 0
                    p = a + b - c
                    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                    pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                r[i] = (sv[i] + pred) & 255
            row = np.array(r, np.int32)
        else:
            raise ValueError(f"{path}: bad PNG filter")
        img[y] = row
        prev = row
    img = img.astype(np.uint8).reshape(H, W, nch)
    if not color:
        if ctype not in (0, 4):
            raise ValueError(f"{path}: a grey PNG is expected")
        return np.ascontiguousarray(img[..., 0])
    if ctype == 3:
        rgb = palette[img[..., 0]]
    elif nch <= 2:
        rgb = np.repeat(img[..., :1], 3, axis=2)
    else:
        rgb = img[..., :3]
    return np.ascontiguousarray(rgb[..., ::-1])


def _encode_png(path, rows, ctype, nch, filters=None, palette=None):
    """Writes raw 8-bit samples rows u8 [H][W*nch] as a PNG of colour type ctype, each row with
    its filter type from filters (default 0); palette u8 [n][3] for ctype 3."""
    rows = np.asarray(rows, np.uint8).astype(np.int32)
    H, n = rows.shape
    W = n // nch
    filters = [0] * H if filters is None else list(filters)
    out = bytearray()
    prev = np.zeros(n, np.int32)
    for y in range(H):
        ft, cur = filters[y], rows[y]
        a = np.concatenate([np.zeros(nch, np.int32), cur[:-nch]])
        c = np.concatenate([np.zeros(nch, np.int32), prev[:-nch]])
        pred = {0: 0, 1: a, 2: prev, 3: (a + prev) >> 1, 4: _paeth(a, prev, c)}[ft]
        out.append(ft)
        out += ((cur - pred) & 255).astype(np.uint8).tobytes()
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, ctype, 0, 0, 0)))
        if ctype == 3:
            f.write(chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes()))
        f.write(chunk(b"IDAT", zlib.compress(bytes(out))))
        f.write(chunk(b"IEND", b""))


def write_png(path, img, filters=None):
    """Writes u8 [H][W] (grey) or [H][W][3] BGR (stored as RGB), each row with its filter type
    from filters (0-4, default 0)."""
    img = np.asarray(img, np.uint8)
    H, W = img.shape[:2]
    if img.ndim == 2:
        _encode_png(path, img.reshape(H, W), 0, 1, filters)
    else:
        _encode_png(path, img[..., ::-1].reshape(H, W * 3), 2, 3, filters)


def error_threshold(input_dir):
    """main_bench_middv3.cu:61-65: 1.0, halved for Q-size scenes, doubled for F-size scenes."""
    if "trainingQ" in input_dir or "testQ" in input_dir:
        return 0.5
    if "trainingF" in input_dir or "testF" in input_dir:
        return 2.0
    return 1.0


def load_scene(input_dir):
    """loadData (main_bench_middv3.cu:12-48): (im0, im1, dispGT, nonocc, calib).  A missing
    disp0GT.pfm gives zeros; a missing mask gives all 255; the mask is (mask == 255) * 255."""
    d = input_dir if input_dir.endswith(os.sep) else input_dir + os.sep
    calib = read_calib(d + "calib.txt")
    if calib is None or calib["ndisp"] <= 0:
        raise ValueError("ndisp is not specified")
    im0, im1 = read_png(d + "im0.png"), read_png(d + "im1.png")
    if im0 is None or im1 is None:
        raise FileNotFoundError(f"im0.png / im1.png not found in {d}")
    gt = read_pfm(d + "disp0GT.pfm")
    if gt is None:
        gt = np.zeros(im0.shape[:2], np.float32)
    mask = read_png(d + "mask0nocc.png", color=False)
    mask = np.full(im0.shape[:2], 255, np.uint8) if mask is None else np.where(mask == 255, 255, 0).astype(np.uint8)
    return im0, im1, gt, mask, calib


def evaluate(ctx, disp, gt, nonocc, threshold):
    """Evaluator::evaluate's three rates (all, nonocc, occ) in percent, counted on the device."""
    c = ctx.eval_bad_pixels(disp, gt, nonocc, threshold)
    return tuple(100.0 * (1.0 - good / n) if n else float("nan") for n, good in c)
