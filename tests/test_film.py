"""Film encoders (include/psy_image.h, csrc/film.cpp) against the oracle's
restatement of the reference's image writers (oracle/psyimage.py), and the
oracle against closed forms.  Host code only."""

import struct
import zlib

import numpy as np
import pytest

from oracle import psyfile as opf
from oracle import psyimage as oim
from psychopath_amd import _lib, film


def sample_film(h, w, seed=0):
    rng = np.random.default_rng(seed)
    img = (rng.random((h, w, 3)) * 1.5).astype(np.float32)
    flat = img.reshape(-1, 3)
    specials = np.array([[0, 0, 0], [1, 1, 1], [np.nan, 0.5, 0.2], [np.inf, -np.inf, 2.0], [-0.3, 1e-9, 0.0031308],
                         [0.0031307, 0.0031309, 0.5], [1e30, -1e30, 65504.0], [7e-8, 3e-5, 6.1e-5]], np.float32)
    flat[:len(specials)] = specials[:len(flat)]
    return img


# ---------------------------------------------------------------- oracle pins

def test_oracle_matrices_are_inverse():
    fwd = opf.rec709_e_matrix().astype(np.float64)
    inv = oim.xyz_to_rec709_e_matrix().astype(np.float64)
    assert np.allclose(fwd @ inv, np.eye(3), atol=3e-7)
    white = oim.xyz_to_rec709_e(np.ones((1, 3), np.float32))[0]
    assert np.allclose(white, 1.0, atol=1e-6)


def test_oracle_srgb_closed_forms():
    n = np.array([0.0, 1.0, 0.0031308, 0.5, 0.18], np.float32)
    g = oim.srgb_gamma(n).astype(np.float64)
    x = n.astype(np.float64)
    ref = np.where(x < 0.0031308, x * 12.92, 1.055 * x ** (1 / 2.4) - 0.055)
    assert np.allclose(g, ref, rtol=1e-6, atol=1e-7)
    assert g[0] == 0.0 and abs(g[1] - 1.0) < 1e-6
    q = oim.quantize(np.array([1.0, 0.5, np.nan, -1.0, 2.0, 0.999, 1 / 255.0], np.float32))
    assert q.tolist() == [255, 127, 0, 0, 255, 254, 1]


# ------------------------------------------------------------ C++ vs oracle

@pytest.mark.parametrize("h,w", [(1, 1), (5, 7), (64, 48), (300, 517)])
def test_ppm_matches_oracle(h, w):
    img = sample_film(h, w)
    assert film.encode(img, _lib.PSY_IMAGE_PPM) == oim.ppm(img, False)
    assert film.encode(img, _lib.PSY_IMAGE_PPM_ASCII) == oim.ppm(img, True)


def png_chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, out = 8, []
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(typ + body) & 0xffffffff
        out.append((typ, body))
        pos += 12 + n
    return out


@pytest.mark.parametrize("h,w", [(1, 1), (5, 7), (64, 48), (300, 517)])
def test_png_matches_oracle(h, w):
    img = sample_film(h, w, 1)
    chunks = png_chunks(film.encode(img, _lib.PSY_IMAGE_PNG))
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    assert struct.unpack(">IIBBBBB", chunks[0][1]) == (w, h, 8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    assert np.array_equal(rows[:, 1:].reshape(h, w, 4), oim.png_pixels(img))


def test_png_decodes_with_an_independent_reader():
    PIL = pytest.importorskip("PIL.Image")
    import io
    img = sample_film(33, 21, 2)
    im = PIL.open(io.BytesIO(film.encode(img, _lib.PSY_IMAGE_PNG)))
    assert im.size == (21, 33) and im.mode == "RGBA"
    assert np.array_equal(np.asarray(im), oim.png_pixels(img))


def read_exr(data):
    assert struct.unpack("<II", data[:8]) == (20000630, 2)
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        typ = data[e + 1:e2].decode()
        n = struct.unpack("<I", data[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, data[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack(f"<{h}Q", data[pos:pos + 8 * h])
    chans = []
    ch = attrs["channels"][1]
    p = 0
    while ch[p] != 0:
        e = ch.index(b"\0", p)
        chans.append((ch[p:e].decode(), struct.unpack("<i", ch[e + 1:e + 5])[0]))
        p = e + 17
    assert chans == [("B", 1), ("G", 1), ("R", 1)]
    assert attrs["compression"][1] == b"\0"
    out = np.zeros((h, w, 3), np.float16)
    for y, off in enumerate(offs):
        yy, n = struct.unpack("<iI", data[off:off + 8])
        assert yy == y and n == 6 * w
        line = np.frombuffer(data[off + 8:off + 8 + n], "<f2").reshape(3, w)
        out[y, :, 2], out[y, :, 1], out[y, :, 0] = line[0], line[1], line[2]
    return out


@pytest.mark.parametrize("h,w", [(1, 1), (5, 7), (64, 48)])
def test_exr_matches_oracle(h, w):
    img = sample_film(h, w, 3)
    got = read_exr(film.encode(img, _lib.PSY_IMAGE_EXR))
    ref = oim.exr_pixels(img)
    both_nan = np.isnan(got) & np.isnan(ref)
    assert np.array_equal(got.view(np.uint16)[~both_nan], ref.view(np.uint16)[~both_nan])


def test_half_rounding_exhaustive_on_boundaries():
    """f32 -> f16 rounding at every half-float boundary and midpoint: the
    library's EXR pixels against numpy's float16 conversion."""
    h16 = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float32)
    mids = ((h16[:-1].astype(np.float64) + h16[1:]) / 2).astype(np.float32)
    vals = np.concatenate([h16, mids, np.nextafter(mids, np.float32(np.inf)), np.nextafter(mids, np.float32(0)),
                           -h16, np.array([70000.0, 65520.0, 65519.0, 2.0 ** -25, 2.0 ** -26], np.float32)])
    # Feed the values straight through: a film whose XYZ maps to them needs
    # the inverse matrix, so check the converter on rec709 = XYZ-mapped
    # values instead by comparing EXR output with the oracle's conversion.
    n = len(vals)
    w = 1024
    h = (n + w - 1) // w
    img = np.zeros((h * w, 3), np.float32)
    img[:n, 0] = vals
    img[:n, 1] = vals[::-1]
    img[:n, 2] = vals * np.float32(0.5)
    img = img.reshape(h, w, 3)
    got = read_exr(film.encode(img, _lib.PSY_IMAGE_EXR))
    ref = oim.exr_pixels(img)
    assert np.array_equal(got.view(np.uint16), ref.view(np.uint16))


@pytest.mark.parametrize("box", [(0, 0, 5, 3), (2, 1, 7, 5), (3, 2, 100, 100), (4, 4, 4, 9), (6, 0, 7, 1)])
def test_bucket_base64_matches_oracle(box):
    img = sample_film(5, 7, 4)
    img = np.nan_to_num(img, nan=0.25, posinf=3.0, neginf=-3.0)
    assert film.bucket_base64(img, *box) == oim.bucket_base64(img, *box)


def test_blender_protocol_lines():
    img = sample_film(40, 70, 5)
    img = np.nan_to_num(img, nan=0.25, posinf=3.0, neginf=-3.0)
    blocks = list(film.blender_buckets(img, bucket=32))
    assert len(blocks) == 2 * 3
    lines = blocks[-1].splitlines()
    assert lines[0] == "DIV" and lines[1] == "100.00%" and lines[2] == "64 32 70 40"
    assert lines[3] == oim.bucket_base64(img, 64, 32, 70, 40) and lines[4:] == ["BUCKET_END", "DIV"]


def test_encode_errors():
    with pytest.raises(_lib.PsyError):
        film.encode(np.zeros((2, 2, 3), np.float32), 9)
    with pytest.raises(ValueError):
        film.write("/tmp/x.jpg", np.zeros((2, 2, 3), np.float32))


def test_rgb_to_xyz_derivation_matches_the_srgb_matrix():
    """The build script's RGBtoXYZ with Rec.709 primaries and D65 white is
    the textbook sRGB matrix (IEC 61966-2-1), pinning the derivation the
    unassigned-shader colour and the equal-energy matrices share."""
    m = np.array(opf._rgb_to_xyz_matrix((0.640, 0.330), (0.300, 0.600), (0.150, 0.060), (0.3127, 0.3290)))
    srgb = np.array([[0.4124, 0.3576, 0.1805], [0.2126, 0.7152, 0.0722], [0.0193, 0.1192, 0.9505]])
    assert np.allclose(m, srgb, atol=1.5e-4)
