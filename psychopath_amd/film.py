"""Film output over the C ABI (include/psy_image.h): the reference's image
writers (Image::write_ascii_ppm, write_binary_ppm, write_png, write_exr,
src/image.rs:101-195) and the Blender bucket protocol (Bucket::rgba_base64,
image.rs:227-252; renderer.rs:312-346) for an XYZ film [H, W, 3]."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib

FORMATS = {".ppm": _lib.PSY_IMAGE_PPM, ".png": _lib.PSY_IMAGE_PNG, ".exr": _lib.PSY_IMAGE_EXR}


def _film(xyz) -> np.ndarray:
    a = np.ascontiguousarray(xyz, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("film must be [H, W, 3]")
    return a


def _call(name, fn, *args):
    n = fn(*args, None, 0)
    if n < 0:
        check(name, int(n))
    buf = C.create_string_buffer(int(n))
    m = fn(*args, buf, int(n))
    if m != n:
        raise RuntimeError(f"{name} changed size")
    return buf.raw[:n]


def encode(xyz, fmt: int) -> bytes:
    """The film as PSY_IMAGE_PPM_ASCII, PSY_IMAGE_PPM, PSY_IMAGE_PNG or
    PSY_IMAGE_EXR file bytes."""
    a = _film(xyz)
    h, w, _ = a.shape
    return _call("psy_image_encode", lib().psy_image_encode, a.ctypes.data_as(_lib._fp), w, h, int(fmt))


def write(path: str, xyz, ascii_ppm: bool = False) -> None:
    """Writes the film to `path`, the format chosen by its extension."""
    ext = os.path.splitext(path)[1].lower()
    if ext not in FORMATS:
        raise ValueError(f"unsupported image extension {ext!r} (use .ppm, .png or .exr)")
    fmt = _lib.PSY_IMAGE_PPM_ASCII if (ext == ".ppm" and ascii_ppm) else FORMATS[ext]
    with open(path, "wb") as f:
        f.write(encode(xyz, fmt))


def bucket_base64(xyz, x0: int, y0: int, x1: int, y1: int) -> str:
    a = _film(xyz)
    h, w, _ = a.shape
    return _call("psy_image_bucket_base64", lib().psy_image_bucket_base64, a.ctypes.data_as(_lib._fp), w, h,
                 x0, y0, x1, y1).decode("ascii")


def blender_buckets(xyz, bucket: int = 32):
    """The reference's serialized output for a finished film: per bucket
    the lines DIV, the progress percentage, `min_x min_y max_x max_y`, the
    base64 payload, BUCKET_END and DIV (renderer.rs:328-336), in scanline
    order of buckets."""
    a = _film(xyz)
    h, w, _ = a.shape
    done, total = 0, w * h
    for y in range(0, h, bucket):
        for x in range(0, w, bucket):
            x1, y1 = min(x + bucket, w), min(y + bucket, h)
            done += (x1 - x) * (y1 - y)
            yield "\n".join(["DIV", "%.2f%%" % (done / total * 100.0), f"{x} {y} {x1} {y1}",
                             bucket_base64(a, x, y, x1, y1), "BUCKET_END", "DIV"]) + "\n"
