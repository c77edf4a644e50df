// psyref — CPU restatement of the psychopath ray-tracing path.
//
// TEST INFRASTRUCTURE ONLY.  This code is the parity oracle and the CPU
// baseline.  Only the test suite, the smoke check and the benchmark's CPU
// baseline leg may load it; the product library never links it.
//
// Math foundation: restates sub_crates/math3d (Point/Vector/Normal/Transform
// over glam 0.15.1 Vec3A/Affine3A on SSE2), src/lerp.rs, src/fp_utils.rs,
// src/hash.rs and src/math.rs.  Build with -ffp-contract=off and without
// fast-math: every expression below is evaluated in the written order with
// IEEE single-precision rounding, as the reference's SSE2 code is.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace psyref {

// SSE2 minps / maxps semantics (glam Vec3A/Vec4 min/max):
// min(a, b) = a < b ? a : b, max(a, b) = a > b ? a : b.  The second operand is
// returned whenever the comparison is false, which includes NaN operands.
static inline float sse_min(float a, float b) { return a < b ? a : b; }
static inline float sse_max(float a, float b) { return a > b ? a : b; }

struct Vec3 {
    float x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

static inline Vec3 operator+(Vec3 a, Vec3 b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline Vec3 operator-(Vec3 a, Vec3 b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline Vec3 operator*(Vec3 a, float s) { return Vec3(a.x * s, a.y * s, a.z * s); }
static inline Vec3 operator*(Vec3 a, Vec3 b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline Vec3 operator/(Vec3 a, float s) { return Vec3(a.x / s, a.y / s, a.z / s); }
// glam Neg flips the sign bit.
static inline Vec3 neg(Vec3 a) { return Vec3(-a.x, -a.y, -a.z); }
static inline Vec3 vmin(Vec3 a, Vec3 b) { return Vec3(sse_min(a.x, b.x), sse_min(a.y, b.y), sse_min(a.z, b.z)); }
static inline Vec3 vmax(Vec3 a, Vec3 b) { return Vec3(sse_max(a.x, b.x), sse_max(a.y, b.y), sse_max(a.z, b.z)); }
// glam dot3: ((x*x') + (y*y')) + (z*z')
static inline float dot(Vec3 a, Vec3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
static inline Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
static inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
static inline float length2(Vec3 a) { return dot(a, a); }
// glam normalize: v * (1 / length)
static inline Vec3 normalized(Vec3 a) { return a * (1.0f / length(a)); }
// glam signum (0.15): x >= 0 -> 1, else -1.
static inline Vec3 signum(Vec3 a) {
    return Vec3(a.x >= 0.0f ? 1.0f : -1.0f, a.y >= 0.0f ? 1.0f : -1.0f, a.z >= 0.0f ? 1.0f : -1.0f);
}
// math3d Vector::abs = co * co.signum()   (vector.rs:44-48)
static inline Vec3 vabs(Vec3 a) { return a * signum(a); }
// glam SSE2 max_element / min_element reduce (x,z) and (y,z) first.
static inline float max_element(Vec3 a) { return sse_max(sse_max(a.x, a.z), sse_max(a.y, a.z)); }
static inline float min_element(Vec3 a) { return sse_min(sse_min(a.x, a.z), sse_min(a.y, a.z)); }
static inline Vec3 recip(Vec3 a) { return Vec3(1.0f / a.x, 1.0f / a.y, 1.0f / a.z); }

// Affine transform = glam Affine3A: three columns plus a translation.
// Transform::new_from_values(a..l) builds columns (a,e,i), (b,f,j), (c,g,k)
// and translation (d,h,l)   (math3d/src/transform.rs:29-47).
struct Transform {
    Vec3 c0, c1, c2, t;
    Transform() : c0(1, 0, 0), c1(0, 1, 0), c2(0, 0, 1), t(0, 0, 0) {}
    static Transform from_values(const float v[12]) {
        Transform m;
        m.c0 = Vec3(v[0], v[4], v[8]);
        m.c1 = Vec3(v[1], v[5], v[9]);
        m.c2 = Vec3(v[2], v[6], v[10]);
        m.t = Vec3(v[3], v[7], v[11]);
        return m;
    }
    void to_values(float v[12]) const {
        v[0] = c0.x; v[4] = c0.y; v[8] = c0.z;
        v[1] = c1.x; v[5] = c1.y; v[9] = c1.z;
        v[2] = c2.x; v[6] = c2.y; v[10] = c2.z;
        v[3] = t.x;  v[7] = t.y;  v[11] = t.z;
    }
};

// glam Mat3A::mul_vec3a without FMA: ((c0*x) + (c1*y)) + (c2*z)
static inline Vec3 mat3_mul(const Transform& m, Vec3 v) {
    Vec3 r = m.c0 * v.x;
    r = (m.c1 * v.y) + r;
    r = (m.c2 * v.z) + r;
    return r;
}
// Point * Transform = transform_point3a: M*p + t   (point.rs:132-141)
static inline Vec3 xform_point(Vec3 p, const Transform& m) { return mat3_mul(m, p) + m.t; }
// Vector * Transform = transform_vector3a: M*v   (vector.rs:141-150)
static inline Vec3 xform_vector(Vec3 v, const Transform& m) { return mat3_mul(m, v); }

// glam Affine3A * Affine3A: (a*b).m = a.m*b.m, (a*b).t = a.m*b.t + a.t.
static inline Transform affine_mul(const Transform& a, const Transform& b) {
    Transform r;
    r.c0 = mat3_mul(a, b.c0);
    r.c1 = mat3_mul(a, b.c1);
    r.c2 = mat3_mul(a, b.c2);
    r.t = mat3_mul(a, b.t) + a.t;
    return r;
}
// math3d `Transform * Transform` applies self first, then other:
// Self(other.0 * self.0)   (transform.rs:89-96)
static inline Transform xform_compose(const Transform& self, const Transform& other) {
    return affine_mul(other, self);
}

// glam Mat3A::inverse: rows of the inverse are the cross products of column
// pairs scaled by 1/det, det = c2 . (c0 x c1).
static inline Transform mat3_inverse(const Transform& m) {
    Vec3 tmp0 = cross(m.c1, m.c2);
    Vec3 tmp1 = cross(m.c2, m.c0);
    Vec3 tmp2 = cross(m.c0, m.c1);
    float det = dot(m.c2, tmp2);
    float inv_det = 1.0f / det;
    Vec3 r0 = tmp0 * inv_det, r1 = tmp1 * inv_det, r2 = tmp2 * inv_det;
    Transform r;
    r.c0 = Vec3(r0.x, r1.x, r2.x);
    r.c1 = Vec3(r0.y, r1.y, r2.y);
    r.c2 = Vec3(r0.z, r1.z, r2.z);
    r.t = Vec3(0, 0, 0);
    return r;
}
// glam Affine3A::inverse: m' = inverse(m), t' = -(m' * t)   (transform.rs:77-79)
static inline Transform xform_inverse(const Transform& a) {
    Transform r = mat3_inverse(a);
    r.t = neg(mat3_mul(r, a.t));
    return r;
}
static inline Transform mat3_transpose(const Transform& m) {
    Transform r;
    r.c0 = Vec3(m.c0.x, m.c1.x, m.c2.x);
    r.c1 = Vec3(m.c0.y, m.c1.y, m.c2.y);
    r.c2 = Vec3(m.c0.z, m.c1.z, m.c2.z);
    r.t = Vec3(0, 0, 0);
    return r;
}
// Normal * Transform = matrix3.inverse().transpose() * n   (normal.rs:129-137)
static inline Vec3 xform_normal(Vec3 n, const Transform& m) {
    return mat3_mul(mat3_transpose(mat3_inverse(m)), n);
}
// Transform * f32 and Transform + Transform are element-wise over the 12
// entries (transform.rs:98-118).
static inline Transform xform_scale(const Transform& a, float s) {
    Transform r;
    r.c0 = a.c0 * s; r.c1 = a.c1 * s; r.c2 = a.c2 * s; r.t = a.t * s;
    return r;
}
static inline Transform xform_add(const Transform& a, const Transform& b) {
    Transform r;
    r.c0 = a.c0 + b.c0; r.c1 = a.c1 + b.c1; r.c2 = a.c2 + b.c2; r.t = a.t + b.t;
    return r;
}

// ---- lerp.rs -------------------------------------------------------------
// Lerp for f32, Vec4, Point, Vector, Normal, Transform: (a*(1-α)) + (b*α)
// (lerp.rs:58-135); the factor (1-α) is rounded once.
static inline float lerp(float a, float b, float alpha) { return (a * (1.0f - alpha)) + (b * alpha); }
static inline Vec3 lerp(Vec3 a, Vec3 b, float alpha) { return (a * (1.0f - alpha)) + (b * alpha); }
static inline Transform lerp(const Transform& a, const Transform& b, float alpha) {
    return xform_add(xform_scale(a, 1.0f - alpha), xform_scale(b, alpha));
}

// lerp_slice (lerp.rs:20-35): last element when len == 1 or alpha == 1.0,
// else i1 = trunc(alpha*(n-1)), alpha2 = alpha*(n-1) - i1.
template <typename T>
static inline T lerp_slice(const T* s, size_t n, float alpha) {
    if (n == 1 || alpha == 1.0f) return s[n - 1];
    float tmp = alpha * (float)(n - 1);
    size_t i1 = (size_t)tmp;
    size_t i2 = i1 + 1;
    float alpha2 = tmp - (float)i1;
    return lerp(s[i1], s[i2], alpha2);
}

// ---- fp_utils.rs ---------------------------------------------------------
// fp_gamma (fp_utils.rs:8-13)
static inline float fp_gamma(uint32_t n) {
    const float e = 1.1920929e-7f * 0.5f;  // f32::EPSILON * 0.5
    return (e * (float)n) / (1.0f - (e * (float)n));
}
static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
// increment_ulp / decrement_ulp (fp_utils.rs:15-48)
static inline float increment_ulp(float v) {
    if (std::isfinite(v)) {
        if (v > 0.0f) return u2f(f2u(v) + 1);
        if (v < -0.0f) return u2f(f2u(v) - 1);
        return u2f(0x00000001u);
    }
    return v;
}
static inline float decrement_ulp(float v) {
    if (std::isfinite(v)) {
        if (v > 0.0f) return u2f(f2u(v) - 1);
        if (v < -0.0f) return u2f(f2u(v) + 1);
        return u2f(0x80000001u);
    }
    return v;
}
// robust_ray_origin (fp_utils.rs:50-82)
static inline Vec3 robust_ray_origin(Vec3 pos, float pos_err, Vec3 nor_in, Vec3 ray_dir) {
    Vec3 nor = dot(nor_in, ray_dir) >= 0.0f ? nor_in : neg(nor_in);
    float d = dot(vabs(nor), Vec3(pos_err, pos_err, pos_err));
    Vec3 offset = nor * d;
    Vec3 p = pos + offset;
    float x = nor.x >= 0.0f ? increment_ulp(p.x) : decrement_ulp(p.x);
    float y = nor.y >= 0.0f ? increment_ulp(p.y) : decrement_ulp(p.y);
    float z = nor.z >= 0.0f ? increment_ulp(p.z) : decrement_ulp(p.z);
    return Vec3(x, y, z);
}

// ---- hash.rs -------------------------------------------------------------
static inline uint32_t hash_u32(uint32_t n, uint32_t seed) {
    uint32_t hash = n;
    for (int i = 0; i < 3; i++) {
        hash = hash * 0x736caf6fu;
        hash ^= hash >> 16;
        hash ^= seed;
    }
    return hash;
}
static inline uint64_t hash_u64(uint64_t n, uint64_t seed) {
    uint64_t hash = n;
    for (int i = 0; i < 4; i++) {
        hash = hash * (32416190071ull * 314604959ull);
        hash ^= hash >> 32;
        hash ^= seed;
    }
    return hash;
}
static inline float hash_u32_to_f32(uint32_t n, uint32_t seed) {
    const float inv_max = 1.0f / (float)4294967295u;
    return (float)hash_u32(n, seed) * inv_max;
}

// ---- math.rs -------------------------------------------------------------
static inline uint64_t log2_64(uint64_t n) {
    if (n == 0) return 0;
    return 63 - (uint64_t)__builtin_clzll(n);
}

}  // namespace psyref
