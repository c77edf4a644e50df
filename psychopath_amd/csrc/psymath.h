// Arithmetic shared by the host builder and the gfx950 kernels.
//
// Every expression keeps the reference's operation order with IEEE single
// precision rounding, so the host packer, the device kernels and the
// reference agree bit for bit.  The reference's vector arithmetic is glam
// 0.15.1 on SSE2: min/max are minps/maxps (second operand on false compares),
// dot products sum left to right, matrix-vector products accumulate column by
// column without FMA.  Compile with -ffp-contract=off and correctly rounded
// division and square root.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PSY_HD __host__ __device__ __forceinline__
#else
#define PSY_HD inline
#endif

namespace psy {

PSY_HD float fmin_sse(float a, float b) { return a < b ? a : b; }
PSY_HD float fmax_sse(float a, float b) { return a > b ? a : b; }

struct V3 {
    float x, y, z;
};
PSY_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
PSY_HD V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PSY_HD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PSY_HD V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
PSY_HD V3 vmin(V3 a, V3 b) { return V3{fmin_sse(a.x, b.x), fmin_sse(a.y, b.y), fmin_sse(a.z, b.z)}; }
PSY_HD V3 vmax(V3 a, V3 b) { return V3{fmax_sse(a.x, b.x), fmax_sse(a.y, b.y), fmax_sse(a.z, b.z)}; }
PSY_HD float dot(V3 a, V3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
PSY_HD V3 cross(V3 a, V3 b) {
    return V3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
PSY_HD float len(V3 a) { return sqrtf(dot(a, a)); }
PSY_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// lerp (src/lerp.rs:58-135): (a*(1-alpha)) + (b*alpha), (1-alpha) rounded once.
PSY_HD float lerpf(float a, float b, float alpha) { return (a * (1.0f - alpha)) + (b * alpha); }
PSY_HD V3 lerp3(V3 a, V3 b, float alpha) {
    float ia = 1.0f - alpha;
    return V3{(a.x * ia) + (b.x * alpha), (a.y * ia) + (b.y * alpha), (a.z * ia) + (b.z * alpha)};
}

// lerp_slice index math (src/lerp.rs:20-35).  Returns false when the last
// sample is to be taken as is (n == 1 or alpha == 1.0); otherwise the segment
// start i1 and local alpha.  The reference requires alpha in [0, 1] (a debug
// assertion, then an out-of-range index); here values outside that range and
// NaN are clamped so the sample index always stays in bounds.
PSY_HD bool lerp_slice_pos(uint32_t n, float alpha, uint32_t& i1, float& a2) {
    if (!(alpha >= 0.0f)) alpha = 0.0f;
    if (alpha > 1.0f) alpha = 1.0f;
    if (n == 1 || alpha == 1.0f) {
        i1 = n - 1;
        a2 = 0.0f;
        return false;
    }
    float tmp = alpha * (float)(n - 1);
    i1 = (uint32_t)tmp;
    a2 = tmp - (float)i1;
    return true;
}

// Affine transform (glam Affine3A): rows m[r][0..2] are the 3x3 part
// (column c = (m[0][c], m[1][c], m[2][c])), m[r][3] the translation.
struct Xf {
    float m[3][4];
};
PSY_HD Xf xf_identity() {
    Xf r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) r.m[i][j] = (i == j) ? 1.0f : 0.0f;
    return r;
}
// From Transform::new_from_values(a..l) (math3d/src/transform.rs:29-47):
// row r is (v[4r], v[4r+1], v[4r+2], v[4r+3]).
PSY_HD Xf xf_from_values(const float* v) {
    Xf r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) r.m[i][j] = v[4 * i + j];
    return r;
}
PSY_HD V3 xf_col(const Xf& a, int c) { return V3{a.m[0][c], a.m[1][c], a.m[2][c]}; }
// Mat3A * v without FMA: ((c0*x) + (c1*y)) + (c2*z), per component.
PSY_HD V3 xf_mat_mul(const Xf& a, V3 v) {
    V3 r;
    r.x = ((a.m[0][0] * v.x) + (a.m[0][1] * v.y)) + (a.m[0][2] * v.z);
    r.y = ((a.m[1][0] * v.x) + (a.m[1][1] * v.y)) + (a.m[1][2] * v.z);
    r.z = ((a.m[2][0] * v.x) + (a.m[2][1] * v.y)) + (a.m[2][2] * v.z);
    return r;
}
// Point * Transform (math3d point.rs:132-141): M*p + t
PSY_HD V3 xf_point(const Xf& a, V3 p) {
    V3 r = xf_mat_mul(a, p);
    return V3{r.x + a.m[0][3], r.y + a.m[1][3], r.z + a.m[2][3]};
}
// Vector * Transform (math3d vector.rs:141-150): M*v
PSY_HD V3 xf_vector(const Xf& a, V3 v) { return xf_mat_mul(a, v); }
// Element-wise lerp of the 12 entries (Transform * f32 + Transform * f32,
// math3d transform.rs:98-118; src/lerp.rs:100-104).
PSY_HD Xf xf_lerp(const Xf& a, const Xf& b, float alpha) {
    Xf r;
    float ia = 1.0f - alpha;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) r.m[i][j] = (a.m[i][j] * ia) + (b.m[i][j] * alpha);
    return r;
}
// math3d `a * b` (transform.rs:89-96): Self(b.0 * a.0), glam's Affine3A
// product, which applies a, then b.
PSY_HD Xf xf_then(const Xf& a, const Xf& b) {
    Xf r;
    for (int j = 0; j < 4; j++) {
        const V3 c = xf_mat_mul(b, V3{a.m[0][j], a.m[1][j], a.m[2][j]});
        r.m[0][j] = c.x;
        r.m[1][j] = c.y;
        r.m[2][j] = c.z;
    }
    r.m[0][3] += b.m[0][3];
    r.m[1][3] += b.m[1][3];
    r.m[2][3] += b.m[2][3];
    return r;
}

// glam Affine3A::inverse: 3x3 inverse from column cross products scaled by
// 1/det (det = c2 . (c0 x c1)), then t' = -(M' * t).
PSY_HD Xf xf_inverse(const Xf& a) {
    V3 c0 = xf_col(a, 0), c1 = xf_col(a, 1), c2 = xf_col(a, 2);
    V3 r0 = cross(c1, c2);
    V3 r1 = cross(c2, c0);
    V3 r2 = cross(c0, c1);
    float det = dot(c2, r2);
    float inv_det = 1.0f / det;
    r0 = mul(r0, inv_det);
    r1 = mul(r1, inv_det);
    r2 = mul(r2, inv_det);
    Xf r;
    r.m[0][0] = r0.x; r.m[0][1] = r0.y; r.m[0][2] = r0.z;
    r.m[1][0] = r1.x; r.m[1][1] = r1.y; r.m[1][2] = r1.z;
    r.m[2][0] = r2.x; r.m[2][1] = r2.y; r.m[2][2] = r2.z;
    r.m[0][3] = 0.0f; r.m[1][3] = 0.0f; r.m[2][3] = 0.0f;
    V3 t = xf_mat_mul(r, V3{a.m[0][3], a.m[1][3], a.m[2][3]});
    r.m[0][3] = -t.x; r.m[1][3] = -t.y; r.m[2][3] = -t.z;
    return r;
}

// Axis-aligned box (src/bbox.rs).
struct Box {
    V3 lo, hi;
};
PSY_HD Box box_empty() {
    const float inf = __builtin_huge_valf();
    return Box{V3{inf, inf, inf}, V3{-inf, -inf, -inf}};
}
PSY_HD Box box_union(const Box& a, const Box& b) { return Box{vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
PSY_HD Box box_union_pt(const Box& a, V3 p) { return Box{vmin(a.lo, p), vmax(a.hi, p)}; }
PSY_HD Box box_lerp(const Box& a, const Box& b, float alpha) {
    return Box{lerp3(a.lo, b.lo, alpha), lerp3(a.hi, b.hi, alpha)};
}
PSY_HD float box_area(const Box& b) {
    V3 d = sub(b.hi, b.lo);
    return (((d.x * d.y) + (d.y * d.z)) + (d.z * d.x)) * 2.0f;
}
PSY_HD float box_diag(const Box& b) { return len(sub(b.hi, b.lo)); }
// BBox::center = min.lerp(max, 0.5) (src/bbox.rs:82-84)
PSY_HD V3 box_center(const Box& b) { return lerp3(b.lo, b.hi, 0.5f); }
// BBox::transformed (src/bbox.rs:52-75): the 8 corners in the reference's order.
PSY_HD Box box_transformed(const Box& b, const Xf& xf) {
    Box r = box_empty();
    for (int i = 0; i < 8; i++) {
        V3 c = V3{(i & 4) ? b.hi.x : b.lo.x, (i & 2) ? b.hi.y : b.lo.y, (i & 1) ? b.hi.z : b.lo.z};
        V3 v = xf_point(xf, c);
        r.lo = vmin(v, r.lo);
        r.hi = vmax(v, r.hi);
    }
    return r;
}

// fp_gamma (src/fp_utils.rs:8-13)
PSY_HD float fp_gamma(uint32_t n) {
    const float e = 1.1920929e-7f * 0.5f;
    return (e * (float)n) / (1.0f - (e * (float)n));
}

constexpr float BBOX_MAXT_ADJUST = 1.00000024f;  // src/bbox4.rs:14

}  // namespace psy
