"""The reference's own unit tests on this path, restated against the oracle.

Known answers come from the reference's test modules:
  src/lerp.rs:141-321, sub_crates/math3d/src/{transform,point,vector,normal}.rs,
  src/fp_utils.rs:84-130, src/algorithm.rs:307-350, src/math.rs:164-187.
"""

import ctypes as C

import numpy as np
import pytest


def fa(*v):
    return np.array(v, np.float32)


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def xf(*v):
    assert len(v) == 12
    return fa(*v)


def xform_mul(L, a, b):
    out = np.zeros(12, np.float32)
    L.psyref_xform_mul(fp(a), fp(b), fp(out))
    return out


def xform_inverse(L, a):
    out = np.zeros(12, np.float32)
    L.psyref_xform_inverse(fp(a), fp(out))
    return out


def xform_lerp(L, a, b, alpha):
    out = np.zeros(12, np.float32)
    L.psyref_xform_lerp(fp(a), fp(b), alpha, fp(out))
    return out


def apply(L, fn, v, m):
    out = np.zeros(3, np.float32)
    getattr(L, fn)(fp(v), fp(m), fp(out))
    return out


def ulps(a, b):
    ia = np.array(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.array(b, np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


@pytest.fixture(scope="module")
def L(oracle):
    return oracle.lib()


# ---- src/lerp.rs:141-321 --------------------------------------------------
@pytest.mark.parametrize("a,b,alpha,want", [(1.0, 2.0, 0.0, 1.0), (1.0, 2.0, 1.0, 2.0), (1.0, 2.0, 0.5, 1.5)])
def test_lerp(L, a, b, alpha, want):
    assert L.psyref_lerp_f32(a, b, alpha) == want


@pytest.mark.parametrize("alpha,want", [(0.0, 0.0), (1.0, 4.0), (0.5, 2.0), (0.25, 1.0), (0.75, 3.0), (0.625, 2.5)])
def test_lerp_slice(L, alpha, want):
    s = fa(0, 1, 2, 3, 4)
    assert L.psyref_lerp_slice_f32(fp(s), 5, alpha) == want


def test_lerp_matrix(L):
    a = xf(0.0, 2.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 11.0)
    b = xf(-1.0, 1.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 11.0, 12.0)
    c1 = xf(-0.25, 1.75, 2.25, 3.25, 4.25, 5.25, 6.25, 7.25, 8.25, 9.25, 10.25, 11.25)
    c2 = xf(-0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5, 11.5)
    c3 = xf(-0.75, 1.25, 2.75, 3.75, 4.75, 5.75, 6.75, 7.75, 8.75, 9.75, 10.75, 11.75)
    np.testing.assert_array_equal(xform_lerp(L, a, b, 0.0), a)
    np.testing.assert_array_equal(xform_lerp(L, a, b, 0.25), c1)
    np.testing.assert_array_equal(xform_lerp(L, a, b, 0.5), c2)
    np.testing.assert_array_equal(xform_lerp(L, a, b, 0.75), c3)
    np.testing.assert_array_equal(xform_lerp(L, a, b, 1.0), b)


@pytest.mark.parametrize("alpha,want", [(0.0, (1.0, 2.0, 1.0)), (1.0, (-2.0, 1.0, -1.0)), (0.5, (-0.5, 1.5, 0.0))])
def test_lerp_point_vector_normal(L, alpha, want):
    out = np.zeros(3, np.float32)
    L.psyref_vec_lerp(fp(fa(1, 2, 1)), fp(fa(-2, 1, -1)), alpha, fp(out))
    np.testing.assert_array_equal(out, fa(*want))


# ---- sub_crates/math3d ----------------------------------------------------
M = xf(1.0, 2.0, 2.0, 1.5, 3.0, 6.0, 7.0, 8.0, 9.0, 2.0, 11.0, 12.0)


def test_transform_multiply(L):  # transform.rs:153-166
    a = M
    b = xf(1.0, 5.0, 9.0, 13.0, 2.0, 6.0, 10.0, 14.0, 3.0, 7.0, 11.0, 15.0)
    c = xf(97.0, 50.0, 136.0, 162.5, 110.0, 60.0, 156.0, 185.0, 123.0, 70.0, 176.0, 207.5)
    np.testing.assert_array_equal(xform_mul(L, a, b), c)


def test_transform_inverse(L):  # transform.rs:168-177 (relative 1e-7)
    a = xf(1.0, 0.33, 0.0, -2.0, 0.0, 1.0, 0.0, 0.0, 2.1, 0.7, 1.3, 0.0)
    prod = xform_mul(L, a, xform_inverse(L, a))
    ident = xf(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)
    # approx::relative_eq with epsilon 1e-7 on each entry
    assert np.all(np.abs(prod - ident) <= 1e-7 * np.maximum(np.abs(prod), np.abs(ident)) + 1e-7)


def test_point_mul_matrix(L):  # point.rs:166-180
    np.testing.assert_array_equal(apply(L, "psyref_point_xform", fa(1.0, 2.5, 4.0), M), fa(15.5, 54.0, 70.0))


def test_point_mul_composition(L):  # point.rs:182-200
    p = fa(1.0, 2.5, 4.0)
    m2 = xf(4.0, 1.0, 2.0, 3.5, 3.0, 6.0, 5.0, 2.0, 2.0, 2.0, 4.0, 12.0)
    pmm1 = apply(L, "psyref_point_xform", p, xform_mul(L, M, m2))
    pmm2 = apply(L, "psyref_point_xform", apply(L, "psyref_point_xform", p, M), m2)
    assert float(np.sum((pmm1 - pmm2) ** 2)) <= 0.00001


def test_vector_mul_matrix(L):  # vector.rs:220-236
    np.testing.assert_array_equal(apply(L, "psyref_vector_xform", fa(1.0, 2.5, 4.0), M), fa(14.0, 46.0, 58.0))


def test_normal_mul_matrix(L):  # normal.rs:209-221 (4 ulps)
    nm = apply(L, "psyref_normal_xform", fa(1.0, 2.5, 4.0), M)
    assert np.all(ulps(nm, fa(-4.0625, 1.78125, -0.03125)) <= 4)


def test_vector_length_normalize_dot_cross(L):  # vector.rs:247-285
    v = fa(1, 2, 3)
    assert abs(L.psyref_vec_length(fp(v)) - 3.7416573867739413) < 1e-6
    n = np.zeros(3, np.float32)
    L.psyref_vec_normalize(fp(v), fp(n))
    np.testing.assert_allclose(n, [0.2672612419124244, 0.5345224838248488, 0.8017837257372732], atol=1e-6)
    assert L.psyref_vec_dot(fp(fa(1, 2, 3)), fp(fa(1.5, 4.5, 2.5))) == 18.0
    c = np.zeros(3, np.float32)
    L.psyref_vec_cross(fp(fa(1, 0, 0)), fp(fa(0, 1, 0)), fp(c))
    np.testing.assert_array_equal(c, fa(0, 0, 1))


# ---- src/fp_utils.rs:84-130 -----------------------------------------------
def test_ulp_stepping(L):
    inc, dec = L.psyref_increment_ulp, L.psyref_decrement_ulp
    assert inc(1.0) > 1.0 and inc(-1.0) > -1.0
    assert dec(1.0) < 1.0 and dec(-1.0) < -1.0
    for z in (0.0, -0.0):
        assert inc(z) > 0.0 and dec(z) < 0.0
    for v in (1.0, -1.0, np.float32(1.2), np.float32(-1.2)):
        v = float(np.float32(v))
        assert dec(inc(v)) == v and inc(dec(v)) == v


def test_fp_gamma_closed_form(L):
    # gamma(n) = n*eps/(1 - n*eps) with eps = 2^-24, rounded in f32
    for n in (2, 3, 5, 7):
        e = np.float32(2.0 ** -24)
        want = np.float32(e * np.float32(n)) / np.float32(np.float32(1.0) - np.float32(e * np.float32(n)))
        assert L.psyref_fp_gamma(n) == want


# ---- src/algorithm.rs:312-350 ---------------------------------------------
@pytest.mark.parametrize("nth", [5, 3, 0, 9])
def test_quick_select(L, nth):
    lst = np.array([8, 9, 7, 4, 6, 1, 0, 5, 3, 2], np.int32)
    L.psyref_quick_select_i32(lst.ctypes.data_as(C.POINTER(C.c_int32)), 10, nth)
    assert lst[nth] == nth
    assert np.all(lst[:nth] < nth) and np.all(lst[nth + 1:] > nth)


# ---- src/math.rs:168-187 --------------------------------------------------
def test_log2_64(L):
    assert L.psyref_log2_64(0) == 0
    for i in range(64):
        assert L.psyref_log2_64(1 << i) == i
    for i in range(8, 64):
        assert L.psyref_log2_64((1 << i) + 227) == i
    for i in range(16, 64):
        assert L.psyref_log2_64((1 << i) + 56369) == i
    for i in range(32, 64):
        assert L.psyref_log2_64((1 << i) + 2514124923) == i


# ---- src/hash.rs (no reference test; checked against a direct restatement
# with Python integers, which catches a wrong constant or shift) --------------
def test_hash_u64_u32(L):
    def h64(n, seed):
        h = n
        for _ in range(4):
            h = (h * (32416190071 * 314604959)) & (2**64 - 1)
            h ^= h >> 32
            h ^= seed
        return h

    def h32(n, seed):
        h = n
        for _ in range(3):
            h = (h * 0x736CAF6F) & 0xFFFFFFFF
            h ^= h >> 16
            h ^= seed
        return h

    rng = np.random.default_rng(7)
    for n, s in rng.integers(0, 2**63, size=(50, 2)):
        assert L.psyref_hash_u64(int(n), int(s)) == h64(int(n), int(s))
        assert L.psyref_hash_u32(int(n) & 0xFFFFFFFF, int(s) & 0xFFFFFFFF) == h32(int(n) & 0xFFFFFFFF,
                                                                                   int(s) & 0xFFFFFFFF)
