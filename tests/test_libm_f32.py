"""The renderer's single-precision sinf, cosf, asinf and acosf
(psychopath_amd/csrc/libm_f32.h) round exactly as the host C library's,
which the reference's f32::sin/cos/asin/acos call: every input in [-1, 1]
(arcs) and with |x| <= 8 (sine, cosine) on a stride (all of them with
PSY_EXHAUSTIVE=1), plus the branch boundaries and special values."""

import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libmathcheck.so")


@pytest.fixture(scope="module")
def mc():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} missing: run __graft_entry__.build()")
    lib = C.CDLL(LIB)
    lib.mc_mismatches.restype = C.c_uint64
    lib.mc_mismatches.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    lib.mc_mismatches_sincos.restype = C.c_uint64
    lib.mc_mismatches_sincos.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    lib.mc_asinf.restype = C.c_float
    lib.mc_asinf.argtypes = [C.c_float]
    lib.mc_acosf.restype = C.c_float
    lib.mc_acosf.argtypes = [C.c_float]
    return lib


def test_matches_host_libm_on_all_of_minus_one_to_one(mc):
    out = (C.c_uint32 * 2)()
    step = 1 if os.environ.get("PSY_EXHAUSTIVE") == "1" else 97
    bad = mc.mc_mismatches(0, 0x3F800000, step, out)
    assert bad == 0, f"first mismatches at bits {out[0]:#x} (asinf), {out[1]:#x} (acosf)"


@pytest.mark.parametrize("lo", [0x31FFFF00, 0x327FFF00, 0x3EFFFF00, 0x3F34FF00, 0x3F799900, 0x3F7FFF00])
def test_branch_boundaries_exactly(mc, lo):
    out = (C.c_uint32 * 2)()
    assert mc.mc_mismatches(lo, lo + 0x200, 1, out) == 0


def test_values():
    lib = C.CDLL(LIB)
    lib.mc_asinf.restype = C.c_float
    lib.mc_asinf.argtypes = [C.c_float]
    lib.mc_acosf.restype = C.c_float
    lib.mc_acosf.argtypes = [C.c_float]
    xs = np.linspace(-1, 1, 1001, dtype=np.float32)
    a = np.array([lib.mc_asinf(float(x)) for x in xs])
    c = np.array([lib.mc_acosf(float(x)) for x in xs])
    assert np.allclose(a, np.arcsin(xs.astype(np.float64)), atol=3e-7)
    assert np.allclose(c, np.arccos(xs.astype(np.float64)), atol=3e-7)
    assert np.isnan(lib.mc_asinf(1.5)) and np.isnan(lib.mc_acosf(-1.5))


def test_sincos_match_host_libm_up_to_8(mc):
    out = (C.c_uint32 * 2)()
    step = 1 if os.environ.get("PSY_EXHAUSTIVE") == "1" else 89
    bad = mc.mc_mismatches_sincos(0, 0x41000000, step, out)
    assert bad == 0, f"first mismatches at bits {out[0]:#x} (sinf), {out[1]:#x} (cosf)"


@pytest.mark.parametrize("lo", [0x397FFF00, 0x3F490F00, 0x3FC90F00, 0x40490F00, 0x4096CB00, 0x40C90F00])
def test_sincos_quadrant_boundaries_exactly(mc, lo):
    out = (C.c_uint32 * 2)()
    assert mc.mc_mismatches_sincos(lo, lo + 0x200, 1, out) == 0
