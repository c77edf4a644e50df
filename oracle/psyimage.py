"""Oracle restatement of the reference's film output (TEST INFRASTRUCTURE
ONLY: imported by tests/ as the checker of the product's encoders).

Follows:
    xyz_to_rec709_e          sub_crates/color/build.rs:102-127 (gjInverse of
                             the equal-energy Rec.709 matrix, build.rs:229-300,
                             printed with ten decimals, read as f32)
    srgb_gamma, xyz_to_srgbe, quantize_tri_255   src/image.rs:282-312
    write_ascii_ppm, write_binary_ppm, write_png pixel order, write_exr
    pixel values             src/image.rs:101-195
    Bucket::rgba_base64      src/image.rs:227-252

Arithmetic is numpy float32 (one IEEE operation at a time, no contraction);
the power goes through the C library's powf (ctypes), as Rust's f32::powf
does: numpy's own float32 power may use a vectorised approximation.
The PNG container bytes of png_encode_mini (not vendored) and OpenEXR's
PIZ compression are not restated: tests compare decoded pixels.
"""

from __future__ import annotations

import base64
import ctypes

import numpy as np

from .psyfile import _rgb_to_xyz_matrix, _round_f32

_libm = ctypes.CDLL("libm.so.6")
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]


def _gj_inverse(m):
    """Imath gjInverse as the colour build script has it (f64)."""
    t = [list(r) for r in m]
    s = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    for i in range(2):
        pivot, size = i, abs(t[i][i])
        for j in range(i + 1, 3):
            if abs(t[j][i]) > size:
                pivot, size = j, abs(t[j][i])
        if pivot != i:
            t[i], t[pivot] = t[pivot], t[i]
            s[i], s[pivot] = s[pivot], s[i]
        for j in range(i + 1, 3):
            f = t[j][i] / t[i][i]
            for k in range(3):
                t[j][k] -= f * t[i][k]
                s[j][k] -= f * s[i][k]
    for i in (2, 1, 0):
        f = t[i][i]
        for j in range(3):
            t[i][j] /= f
            s[i][j] /= f
        for j in range(i):
            g = t[j][i]
            for k in range(3):
                t[j][k] -= g * t[i][k]
                s[j][k] -= g * s[i][k]
    return s


def xyz_to_rec709_e_matrix() -> np.ndarray:
    m = _rgb_to_xyz_matrix((0.640, 0.330), (0.300, 0.600), (0.150, 0.060), (1.0 / 3.0, 1.0 / 3.0))
    inv = _gj_inverse(m)
    return np.array([[_round_f32("%.10f" % v) for v in row] for row in inv], np.float32)


def xyz_to_rec709_e(xyz: np.ndarray) -> np.ndarray:
    """[..., 3] float32 -> [..., 3] float32: (x*m0 + y*m1) + z*m2 per row."""
    m = xyz_to_rec709_e_matrix()
    x, y, z = (np.asarray(xyz, np.float32)[..., k] for k in range(3))
    return np.stack([(x * m[i, 0] + y * m[i, 1]) + z * m[i, 2] for i in range(3)], axis=-1)


def powf(x: np.ndarray, y: float) -> np.ndarray:
    x = np.asarray(x, np.float32)
    out = np.empty_like(x)
    flat, of = x.ravel(), out.ravel()
    yy = float(np.float32(y))
    for i in range(flat.size):
        of[i] = _libm.powf(float(flat[i]), yy)
    return out


def srgb_gamma(n: np.ndarray) -> np.ndarray:
    n = np.asarray(n, np.float32)
    lo = n < np.float32(0.0031308)
    hi = np.zeros_like(n)
    sel = ~lo
    p = powf(n[sel], np.float32(1.0) / np.float32(2.4))
    hi[sel] = (np.float32(1.055) * p) - np.float32(0.055)
    return np.where(lo, n * np.float32(12.92), hi).astype(np.float32)


def quantize(n: np.ndarray) -> np.ndarray:
    """1.0.min(0.0.max(n)) * 255 as u8 (NaN -> 0)."""
    n = np.asarray(n, np.float32)
    c = np.where(n > 0, n, np.float32(0.0))
    c = np.where(c < 1, c, np.float32(1.0))
    return (c * np.float32(255.0)).astype(np.uint8)


def srgb8(xyz: np.ndarray) -> np.ndarray:
    return quantize(srgb_gamma(xyz_to_rec709_e(xyz)))


def ppm(xyz: np.ndarray, ascii_ppm: bool) -> bytes:
    h, w, _ = xyz.shape
    px = srgb8(xyz)
    head = f"{'P3' if ascii_ppm else 'P6'}\n{w} {h}\n255\n".encode()
    if not ascii_ppm:
        return head + px.tobytes()
    lines = []
    for y in range(h):
        lines.append("".join(f"{p[0]} {p[1]} {p[2]} " for p in px[y]) + "\n")
    return head + "".join(lines).encode()


def png_pixels(xyz: np.ndarray) -> np.ndarray:
    """RGBA rows as write_png orders them: the film's last row first."""
    px = srgb8(xyz)[::-1]
    alpha = np.full(px.shape[:2] + (1,), 255, np.uint8)
    return np.concatenate([px, alpha], axis=-1)


def exr_pixels(xyz: np.ndarray) -> np.ndarray:
    """Half-float R, G, B per pixel (f16::from_f32 rounds to nearest even)."""
    return xyz_to_rec709_e(xyz).astype(np.float16)


def bucket_base64(xyz: np.ndarray, x0, y0, x1, y1) -> str:
    h, w, _ = xyz.shape
    x1, y1 = min(x1, w), min(y1, h)
    if not (x0 < x1 and y0 < y1):
        return ""
    rgb = xyz_to_rec709_e(xyz[y0:y1, x0:x1])
    rgba = np.concatenate([rgb, np.ones(rgb.shape[:2] + (1,), np.float32)], axis=-1)
    return base64.b64encode(rgba.astype("<f4").tobytes()).decode("ascii")
