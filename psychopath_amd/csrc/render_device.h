// Device-side path-tracing math for gfx950: sample streams, colour science,
// lights, surface closures and the light tree.  Every function keeps the
// reference's operation order and precision (f32, with f64 where the
// reference computes in f64), so a path computes the same values as the
// reference's CPU path.  Single-precision sin, cos, asin and acos follow the
// host library's rounding (libm_f32.h), because the light tree's estimates
// and glossy reflections amplify last-bit differences; the double-precision
// functions are the device library's.
//
// The header also compiles as plain host C++ (PSY_RD_HOST defined), which is
// how the tests compare these functions with the oracle's on the CPU.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/psy_render.h"
#include "layout.h"
#include "libm_f32.h"
#include "psymath.h"

#ifdef PSY_RD_HOST
#define PSY_MENG_STORAGE static
#define RD static inline
#else
#include <hip/hip_runtime.h>
#define PSY_MENG_STORAGE static __device__
#define RD __device__ __forceinline__
#endif
#include "meng_tables.h"

namespace psy {
namespace rd {

constexpr float PI_F = 3.14159265358979323846264338327950288f;
constexpr double PI_D = 3.14159265358979323846264338327950288;
constexpr float QPI_F = 0.785398163397448309615660845819875721f;
constexpr float INV_PI_F = 1.0f / PI_F;  // surface_closure.rs:14
constexpr float H_PI_F = PI_F / 2.0f;    // surface_closure.rs:15

// ---- small vector helpers ---------------------------------------------------

// A hero-wavelength spectral sample: four energies (SpectralSample,
// src/color.rs:283-300).
struct S4 {
    float v[4];
};
RD S4 s4(float a, float b, float c, float d) { return S4{{a, b, c, d}}; }
RD S4 s4_splat(float a) { return S4{{a, a, a, a}}; }
RD S4 s4_zero() { return S4{{0.0f, 0.0f, 0.0f, 0.0f}}; }
RD S4 operator+(S4 a, S4 b) { return s4(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]); }
RD S4 operator-(S4 a, S4 b) { return s4(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]); }
RD S4 operator*(S4 a, S4 b) { return s4(a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]); }
RD S4 operator*(S4 a, float s) { return s4(a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s); }
RD S4 operator/(S4 a, float s) { return s4(a.v[0] / s, a.v[1] / s, a.v[2] / s, a.v[3] / s); }
// glam Vec4::max_element on SSE2: max(max(x, z), max(y, w)).
RD float max_element4(S4 a) {
    float m0 = a.v[0] > a.v[2] ? a.v[0] : a.v[2];
    float m1 = a.v[1] > a.v[3] ? a.v[1] : a.v[3];
    return m0 > m1 ? m0 : m1;
}
#ifndef PSY_RD_HOST
RD S4 s4_load(const float4* p, uint32_t i) {
    float4 f = p[i];
    return s4(f.x, f.y, f.z, f.w);
}
RD void s4_store(float4* p, uint32_t i, S4 a) { p[i] = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]); }
#endif

RD V3 neg3(V3 a) { return v3(-a.x, -a.y, -a.z); }
RD float length2(V3 a) { return dot(a, a); }
// glam normalize: v * (1 / length)
RD V3 normalize(V3 a) { return mul(a, 1.0f / len(a)); }
// math3d Vector::abs = co * co.signum(), glam 0.15 signum: x >= 0 -> 1.
RD V3 vabs_sig(V3 a) {
    return v3(a.x * (a.x >= 0.0f ? 1.0f : -1.0f), a.y * (a.y >= 0.0f ? 1.0f : -1.0f),
              a.z * (a.z >= 0.0f ? 1.0f : -1.0f));
}
// glam SSE2 max_element: reduces (x, z) and (y, z) first.
RD float max_element3(V3 a) { return fmax_sse(fmax_sse(a.x, a.z), fmax_sse(a.y, a.z)); }
RD V3 lerp_v3(V3 a, V3 b, float alpha) { return lerp3(a, b, alpha); }

// Normal * Transform: inverse(3x3)^T * n (math3d normal.rs:129-137).
RD V3 xf_normal(const Xf& m, V3 n) {
    V3 c0 = xf_col(m, 0), c1 = xf_col(m, 1), c2 = xf_col(m, 2);
    V3 t0 = cross(c1, c2), t1 = cross(c2, c0), t2 = cross(c0, c1);
    float det = dot(c2, t2);
    float inv_det = 1.0f / det;
    V3 r0 = mul(t0, inv_det), r1 = mul(t1, inv_det), r2 = mul(t2, inv_det);
    // rows of the inverse are r0, r1, r2; the transpose's columns are r0, r1, r2
    return v3(((r0.x * n.x) + (r1.x * n.y)) + (r2.x * n.z), ((r0.y * n.x) + (r1.y * n.y)) + (r2.y * n.z),
              ((r0.z * n.x) + (r1.z * n.y)) + (r2.z * n.z));
}

// increment_ulp / decrement_ulp (src/fp_utils.rs:15-48)
RD float inc_ulp(float v) {
    if (isfinite(v)) {
        if (v > 0.0f) return lm_float(lm_bits(v) + 1u);
        if (v < -0.0f) return lm_float(lm_bits(v) - 1u);
        return lm_float(0x00000001u);
    }
    return v;
}
RD float dec_ulp(float v) {
    if (isfinite(v)) {
        if (v > 0.0f) return lm_float(lm_bits(v) - 1u);
        if (v < -0.0f) return lm_float(lm_bits(v) + 1u);
        return lm_float(0x80000001u);
    }
    return v;
}
// robust_ray_origin (src/fp_utils.rs:50-82)
RD V3 robust_ray_origin(V3 pos, float pos_err, V3 nor_in, V3 ray_dir) {
    V3 nor = dot(nor_in, ray_dir) >= 0.0f ? nor_in : neg3(nor_in);
    float d = dot(vabs_sig(nor), v3(pos_err, pos_err, pos_err));
    V3 p = add(pos, mul(nor, d));
    return v3(nor.x >= 0.0f ? inc_ulp(p.x) : dec_ulp(p.x), nor.y >= 0.0f ? inc_ulp(p.y) : dec_ulp(p.y),
              nor.z >= 0.0f ? inc_ulp(p.z) : dec_ulp(p.z));
}

// surface_point (src/surface/triangle.rs:151-165)
RD void surface_point(V3 t0, V3 t1, V3 t2, float b0, float b1, float b2, V3& pos, float& pos_err) {
    pos = add(add(mul(t0, b0), mul(t1, b1)), mul(t2, b2));
    V3 e = add(add(mul(vabs_sig(t0), b0), mul(vabs_sig(t1), b1)), mul(vabs_sig(t2), b2));
    pos_err = max_element3(mul(e, fp_gamma(7)));
}

// ---- hashing and sample streams ------------------------------------------------

// hash_u32 (src/hash.rs:1-10)
RD uint32_t hash_u32(uint32_t n, uint32_t seed) {
    uint32_t h = n;
    for (int i = 0; i < 3; i++) {
        h = h * 0x736caf6fu;
        h ^= h >> 16;
        h ^= seed;
    }
    return h;
}
// hash_u32_to_f32 (src/hash.rs:26-29)
RD float hash_u32_to_f32(uint32_t n, uint32_t seed) {
    const float inv_max = 1.0f / 4294967296.0f;  // 1 / (u32::MAX as f32)
    return (float)hash_u32(n, seed) * inv_max;
}

// One dimension of the Faure-permuted Halton sequence
// (sub_crates/halton/build.rs): the digit permutation of `pow_base` =
// base^k <= 500 digits at once, applied to successive digit groups.
struct HaltonDim {
    uint32_t base, pow_base, max_power, perm_off;
};
RD float halton_sample(const HaltonDim* dims, const uint16_t* perms, uint32_t dimension, uint32_t index) {
    if (dimension == 0) {
        // Base 2: the bit-reversed index (build.rs:110-126).
#ifdef PSY_RD_HOST
        index = (index << 16) | (index >> 16);
        index = ((index & 0x00ff00ffu) << 8) | ((index & 0xff00ff00u) >> 8);
        index = ((index & 0x0f0f0f0fu) << 4) | ((index & 0xf0f0f0f0u) >> 4);
        index = ((index & 0x33333333u) << 2) | ((index & 0xccccccccu) >> 2);
        index = ((index & 0x55555555u) << 1) | ((index & 0xaaaaaaaau) >> 1);
        return (float)index * (1.0f / 4294967296.0f);
#else
        return (float)__builtin_bitreverse32(index) * (1.0f / 4294967296.0f);
#endif
    }
    const HaltonDim d = dims[dimension];
    const uint16_t* perm = perms + d.perm_off;
    uint32_t power = d.max_power / d.pow_base;
    uint32_t sum = (uint32_t)perm[index % d.pow_base] * power;
    uint32_t div = 1;
    while (power / d.pow_base > 1) {
        div *= d.pow_base;
        power /= d.pow_base;
        sum += (uint32_t)perm[(index / div) % d.pow_base] * power;
    }
    sum += (uint32_t)perm[(index / (div * d.pow_base)) % d.pow_base];
    return (float)sum * (0.999999940395355224609375f / (float)d.max_power);
}

// get_sample_4d (src/renderer.rs:698-725) with the Halton sequence in place
// of sobol_burley: dimension sets below 32 take Halton dimensions 4*set+k at
// index `i`, each rotated modulo 1 by hash_u32_to_f32(dim, pixel seed);
// higher sets use the reference's hash fallback.
RD S4 sample_4d(const HaltonDim* dims, const uint16_t* perms, uint32_t i, uint32_t set, uint32_t px, uint32_t py,
                uint32_t seed) {
    const uint32_t s = px ^ (py << 16) ^ (seed * 0x736caf6fu);
    S4 out;
    if (set < 32u) {
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t dim = set * 4u + k;
            float v = halton_sample(dims, perms, dim, i) + hash_u32_to_f32(dim, s);
            if (v >= 1.0f) v = v - 1.0f;
            out.v[k] = v;
        }
    } else {
        for (uint32_t k = 0; k < 4; k++) out.v[k] = hash_u32_to_f32((set * 4u + k) ^ (i << 16), s);
    }
    return out;
}

// golden_ratio_sample (src/renderer.rs:728-737)
RD float golden_ratio_sample(uint32_t i, uint32_t scramble) {
    uint32_t n = (i + hash_u32(scramble, 0)) * 2654435769u;
    return (float)n * (1.0f / 4294967296.0f);
}

// probit (src/math.rs:67-132): Acklam's inverse normal CDF in f64.
RD float probit(float n_f, float width) {
    const double n = (double)n_f;
    const double A1 = -3.969683028665376e+01, A2 = 2.209460984245205e+02, A3 = -2.759285104469687e+02,
                 A4 = 1.383577518672690e+02, A5 = -3.066479806614716e+01, A6 = 2.506628277459239e+00;
    const double B1 = -5.447609879822406e+01, B2 = 1.615858368580409e+02, B3 = -1.556989798598866e+02,
                 B4 = 6.680131188771972e+01, B5 = -1.328068155288572e+01;
    const double C1 = -7.784894002430293e-03, C2 = -3.223964580411365e-01, C3 = -2.400758277161838e+00,
                 C4 = -2.549732539343734e+00, C5 = 4.374664141464968e+00, C6 = 2.938163982698783e+00;
    const double D1 = 7.784695709041462e-03, D2 = 3.224671290700398e-01, D3 = 2.445134137142996e+00,
                 D4 = 3.754408661907416e+00;
    const double LOW = 0.02425, HIGH = 1.0 - LOW;
    double x;
    if (0.0 < n && n < LOW) {
        const double q = sqrt(-2.0 * log(n));
        x = (((((C1 * q + C2) * q + C3) * q + C4) * q + C5) * q + C6) / ((((D1 * q + D2) * q + D3) * q + D4) * q + 1.0);
    } else if (n <= HIGH) {
        const double q = n - 0.5;
        const double r = q * q;
        x = (((((A1 * r + A2) * r + A3) * r + A4) * r + A5) * r + A6) * q /
            (((((B1 * r + B2) * r + B3) * r + B4) * r + B5) * r + 1.0);
    } else if (n < 1.0) {
        const double q = sqrt(-2.0 * log(1.0 - n));
        x = -(((((C1 * q + C2) * q + C3) * q + C4) * q + C5) * q + C6) /
            ((((D1 * q + D2) * q + D3) * q + D4) * q + 1.0);
    } else if (n == 0.0) {
        x = -__builtin_huge_val();
    } else if (n == 1.0) {
        x = __builtin_huge_val();
    } else {
        x = __builtin_nan("");
    }
    return (float)x * width;
}

// ---- Monte Carlo mappings (src/sampling/monte_carlo.rs) -------------------------

// square_to_circle (monte_carlo.rs:10-32): Shirley's concentric mapping.
RD void square_to_circle(float x, float y, float& ox, float& oy) {
    if (x == 0.0f && y == 0.0f) {
        ox = 0.0f;
        oy = 0.0f;
        return;
    }
    float radius, angle;
    if (x > fabsf(y)) {
        radius = x;
        angle = QPI_F * (y / x);
    } else if (y > fabsf(x)) {
        radius = y;
        angle = QPI_F * (2.0f - (x / y));
    } else if (x < -fabsf(y)) {
        radius = -x;
        angle = QPI_F * (4.0f + (y / x));
    } else {
        radius = -y;
        angle = QPI_F * (6.0f - (x / y));
    }
    ox = radius * psy_cosf(angle);
    oy = radius * psy_sinf(angle);
}
// cosine_sample_hemisphere (monte_carlo.rs:34-38)
RD V3 cosine_sample_hemisphere(float u, float v) {
    float uu, vv;
    square_to_circle((u * 2.0f) - 1.0f, (v * 2.0f) - 1.0f, uu, vv);
    float t = 1.0f - ((uu * uu) + (vv * vv));
    return v3(uu, vv, sqrtf(t > 0.0f ? t : 0.0f));
}
// uniform_sample_sphere (monte_carlo.rs:49-56)
RD V3 uniform_sample_sphere(float u, float v) {
    float z = 1.0f - (2.0f * u);
    float t = 1.0f - (z * z);
    float r = sqrtf(t > 0.0f ? t : 0.0f);
    float phi = (2.0f * PI_F) * v;
    return v3(r * psy_cosf(phi), r * psy_sinf(phi), z);
}
// uniform_sample_cone (monte_carlo.rs:64-73), in f64
RD V3 uniform_sample_cone(float u, float v, double cos_theta_max) {
    double cos_theta = (1.0 - (double)u) + ((double)u * cos_theta_max);
    double sin_theta = sqrt(1.0 - (cos_theta * cos_theta));
    double phi = ((double)v * 2.0) * PI_D;
    return v3((float)(cos(phi) * sin_theta), (float)(sin(phi) * sin_theta), (float)cos_theta);
}
RD double uniform_sample_cone_pdf(double cos_theta_max) { return 1.0 / ((2.0 * PI_D) * (1.0 - cos_theta_max)); }
// uniform_sample_triangle (monte_carlo.rs:82-89)
RD V3 uniform_sample_triangle(V3 va, V3 vb, V3 vc, float i, float j) {
    float isqrt = sqrtf(i);
    float a = 1.0f - isqrt;
    float b = isqrt * (1.0f - j);
    float c = j * isqrt;
    return add(add(mul(va, a), mul(vb, b)), mul(vc, c));
}
// triangle_surface_area (monte_carlo.rs:92-94)
RD float triangle_surface_area(V3 p0, V3 p1, V3 p2) { return 0.5f * len(cross(sub(p1, p0), sub(p2, p0))); }
// f64::max(-1).min(1) and f32::max(-1).min(1): NaN -> -1.
RD double clamp11(double x) { return (x != x) ? -1.0 : (x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x)); }
RD float clamp11f(float x) {
    float m = (x != x) ? -1.0f : (x > -1.0f ? x : -1.0f);
    return m < 1.0f ? m : 1.0f;
}
// spherical_triangle_solid_angle (monte_carlo.rs:99-131)
RD float spherical_triangle_solid_angle(V3 va, V3 vb, V3 vc) {
    double cos_a = (double)clamp11f(dot(vb, vc));
    double cos_b = (double)clamp11f(dot(vc, va));
    double cos_c = (double)clamp11f(dot(va, vb));
    double sin_a = sqrt(1.0 - (cos_a * cos_a));
    double sin_b = sqrt(1.0 - (cos_b * cos_b));
    double sin_c = sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return 0.0f;
    double cos_va = clamp11((cos_a - (cos_b * cos_c)) / (sin_b * sin_c));
    double cos_vb = clamp11((cos_b - (cos_c * cos_a)) / (sin_c * sin_a));
    double cos_vc = clamp11((cos_c - (cos_a * cos_b)) / (sin_a * sin_b));
    return (float)(((acos(cos_va) + acos(cos_vb)) + acos(cos_vc)) - PI_D);
}
// uniform_sample_spherical_triangle (monte_carlo.rs:135-200), Arvo's method
RD V3 uniform_sample_spherical_triangle(V3 va, V3 vb, V3 vc, float i, float j) {
    double cos_a = (double)clamp11f(dot(vb, vc));
    double cos_b = (double)clamp11f(dot(vc, va));
    double cos_c = (double)clamp11f(dot(va, vb));
    double sin_a = sqrt(1.0 - (cos_a * cos_a));
    double sin_b = sqrt(1.0 - (cos_b * cos_b));
    double sin_c = sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return va;
    double cos_va = clamp11((cos_a - (cos_b * cos_c)) / (sin_b * sin_c));
    double cos_vb = clamp11((cos_b - (cos_c * cos_a)) / (sin_c * sin_a));
    double cos_vc = clamp11((cos_c - (cos_a * cos_b)) / (sin_a * sin_b));
    double sin_va = sqrt(1.0 - (cos_va * cos_va));
    double ang_va = acos(cos_va), ang_vb = acos(cos_vb), ang_vc = acos(cos_vc);
    double area = ((ang_va + ang_vb) + ang_vc) - PI_D;
    double area_2 = area * (double)i;
    double s = sin(area_2 - ang_va);
    double t = cos(area_2 - ang_va);
    double u = t - cos_va;
    double v = s + (sin_va * cos_c);
    double q_top = (((v * t) - (u * s)) * cos_va) - v;
    double q_bottom = ((v * s) + (u * t)) * sin_va;
    double q = q_top / q_bottom;
    V3 vc_2 = add(mul(va, (float)q), mul(normalize(sub(vc, mul(va, dot(vc, va)))), (float)sqrt(1.0 - (q * q))));
    float z = 1.0f - (j * (1.0f - dot(vc_2, vb)));
    return add(mul(vb, z), mul(normalize(sub(vc_2, mul(vb, dot(vc_2, vb)))), sqrtf(1.0f - (z * z))));
}

// coordinate_system_from_vector and zup_to_vec (src/math.rs:19-54)
RD void coordinate_system_from_vector(V3 v, V3& v1, V3& v2, V3& v3o) {
    float sign = signbit(v.z) ? -1.0f : 1.0f;  // f32::signum
    float a = -1.0f / (sign + v.z);
    float b = (v.x * v.y) * a;
    v1 = v;
    v2 = v3(1.0f + (((sign * v.x) * v.x) * a), sign * b, -sign * v.x);
    v3o = v3(b, sign + ((v.y * v.y) * a), -v.y);
}
RD V3 zup_to_vec(V3 from, V3 toz) {
    V3 tz, tx, ty;
    coordinate_system_from_vector(normalize(toz), tz, tx, ty);
    return add(add(mul(tx, from.x), mul(ty, from.y)), mul(tz, from.z));
}

// ---- colour (src/color.rs) ------------------------------------------------------

// fastapprox 0.3 fastpow2 / fastexp (Mineiro), with Rust's saturating casts.
RD uint32_t f32_to_u32_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
RD int32_t f32_to_i32_sat(float f) {
    if (f != f) return 0;
    if (f <= -2147483648.0f) return (int32_t)0x80000000u;
    if (f >= 2147483648.0f) return 0x7fffffff;
    return (int32_t)f;
}
RD float fast_pow2(float p) {
    float offset = p < 0.0f ? 1.0f : 0.0f;
    float clipp = p < -126.0f ? -126.0f : p;
    int32_t w = f32_to_i32_sat(clipp);
    float z = (clipp - (float)w) + offset;
    float v = (float)(1 << 23) * (((clipp + 121.2740575f) + (27.7280233f / (4.84252568f - z))) - (1.49012907f * z));
    return lm_float(f32_to_u32_sat(v));
}
RD float fast_exp(float p) { return fast_pow2(1.442695040f * p); }

constexpr float WL_MIN = 380.0f, WL_MAX = 700.0f;
constexpr float WL_RANGE = WL_MAX - WL_MIN;
constexpr float WL_RANGE_Q = WL_RANGE / 4.0f;
RD float map_0_1_to_wavelength(float n) { return (n * WL_RANGE) + WL_MIN; }
RD float nth_wavelength(float hero, int n) {
    float wl = hero + (WL_RANGE_Q * (float)n);
    return wl > WL_MAX ? wl - WL_RANGE : wl;
}

// Meng et al.'s upsampling (sub_crates/spectral_upsampling/src/meng.rs:
// 177-328): xy -> uv grid cell, then bilinear (inside cells) or barycentric
// (border cells) interpolation of tabulated spectra at four wavelengths.
RD S4 spectrum_xyz_to_p_4(S4 lambdas, float x, float y, float z) {
    using namespace psy_meng;
    float inv_norm = (x + y) + z;
    float norm = 1.0f / inv_norm;
    if (!(norm < 3.40282347e+38f)) return s4_zero();
    float xx = x * norm, yy = y * norm;
    float u = ((MAT_XY_TO_UV[0] * xx) + (MAT_XY_TO_UV[1] * yy)) + MAT_XY_TO_UV[2];
    float v = ((MAT_XY_TO_UV[3] * xx) + (MAT_XY_TO_UV[4] * yy)) + MAT_XY_TO_UV[5];
    if (u < 0.0f || u >= (float)GRID_WIDTH || v < 0.0f || v >= (float)GRID_HEIGHT) return s4_zero();
    const int ui = (int)u, vi = (int)v;
    const int* cell = GRID[ui + GRID_WIDTH * vi];
    const bool inside = cell[0] != 0;
    const int num = cell[1];
    const int* idx = cell + 2;
    if (num == 0) return s4_zero();

    S4 sb;
    int sb0[4], sb1[4];
    for (int k = 0; k < 4; k++) {
        sb.v[k] = ((lambdas.v[k] - SAMPLE_MIN) / (SAMPLE_MAX - SAMPLE_MIN)) * ((float)NUM_SAMPLES - 1.0f);
        sb0[k] = (int)sb.v[k];
        sb1[k] = sb0[k] + 1 < NUM_SAMPLES - 1 ? sb0[k] + 1 : NUM_SAMPLES - 1;
    }
    S4 sbf = sb - s4((float)sb0[0], (float)sb0[1], (float)sb0[2], (float)sb0[3]);
    // A cell has at most six points; the loops are unrolled so the points
    // are indexed by constants and stay in registers.
    S4 p[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        p[i] = s4_zero();
        if (i < num) {
            const float* spectrum = POINT_SPECTRUM[idx[i]];
            S4 p0 = s4(spectrum[sb0[0]], spectrum[sb0[1]], spectrum[sb0[2]], spectrum[sb0[3]]);
            S4 p1 = s4(spectrum[sb1[0]], spectrum[sb1[1]], spectrum[sb1[2]], spectrum[sb1[3]]);
            p[i] = (p0 * (s4_splat(1.0f) - sbf)) + (p1 * sbf);
        }
    }
    S4 interp = s4_zero();
    if (inside) {
        float u2 = u - (float)ui, v2 = v - (float)vi;
        interp = (((p[0] * ((1.0f - u2) * (1.0f - v2))) + (p[2] * ((1.0f - u2) * v2))) + (p[3] * (u2 * v2))) +
                 (p[1] * (u2 * (1.0f - v2)));
    } else {
        float ex = u - POINT_UV[idx[0]][0];
        float ey = v - POINT_UV[idx[0]][1];
        float e0x = POINT_UV[idx[1]][0] - POINT_UV[idx[0]][0];
        float e0y = POINT_UV[idx[1]][1] - POINT_UV[idx[0]][1];
        float uu = (e0x * ey) - (ex * e0y);
        bool found = false;
#pragma unroll
        for (int i = 0; i < 5; i++) {
            if (found || i >= num - 1) continue;
            float e1x, e1y;
            if (i == num - 2) {
                e1x = POINT_UV[idx[1]][0] - POINT_UV[idx[0]][0];
                e1y = POINT_UV[idx[1]][1] - POINT_UV[idx[0]][1];
            } else {
                e1x = POINT_UV[idx[i + 2]][0] - POINT_UV[idx[0]][0];
                e1y = POINT_UV[idx[i + 2]][1] - POINT_UV[idx[0]][1];
            }
            float vv = (ex * e1y) - (e1x * ey);
            float area = (e0x * e1y) - (e1x * e0y);
            float uf = uu / area;
            float vf = vv / area;
            float wf = (1.0f - uf) - vf;
            if (uf < 0.0f || vf < 0.0f || wf < 0.0f) {
                uu = -vv;
                e0x = e1x;
                e0y = e1y;
                continue;
            }
            // i + 2 < num here unless i == num - 2, so the clamp never binds.
            const S4 pc = i == num - 2 ? p[1] : p[i + 2 < 6 ? i + 2 : 5];
            interp = ((p[0] * wf) + (p[i + 1] * vf)) + (pc * uf);
            found = true;
        }
    }
    return interp * inv_norm;
}

// plancks_law (color.rs:335-376), per micrometre
RD float plancks_law(float temperature, float wavelength) {
    const float C = 299792458.0f;
    const float H = 6.62607015e-34f;
    const float KB = 1.38064852e-23f;
    if (temperature < 200.0f) return 0.0f;
    float wl = wavelength * 1.0e-9f;
    const float TMP1 = (float)(2.0 * (double)H * (double)C * (double)C);
    const float TMP2 = (float)((double)H * (double)C / (double)KB);
    float wl2 = wl * wl;
    float wl5 = (wl2 * wl2) * wl;
    float tmp3 = wl5 * (fast_exp(TMP2 / (wl * temperature)) - 1.0f);
    float energy = TMP1 / tmp3;
    float e = energy * 1.0e-6f;
    return e > 0.0f ? e : 0.0f;
}
RD float plancks_law_normalized(float temperature, float wavelength) {
    float t2 = temperature * temperature;
    return (plancks_law(temperature, wavelength) * 4.0e7f) / (t2 * t2);
}

// Color::to_spectral_sample (color.rs:64-108) at the four hero wavelengths.
RD S4 color_spectral(const GpuColor& c, float hero) {
    S4 wls = s4(nth_wavelength(hero, 0), nth_wavelength(hero, 1), nth_wavelength(hero, 2), nth_wavelength(hero, 3));
    if (c.kind == PSY_COLOR_XYZ)
        return spectrum_xyz_to_p_4(wls, c.a, c.b, c.c) * s4_splat(1.0f / psy_meng::EQUAL_ENERGY_REFLECTANCE);
    S4 r;
    for (int k = 0; k < 4; k++)
        r.v[k] = (c.kind == PSY_COLOR_BLACKBODY ? plancks_law(c.a, wls.v[k]) : plancks_law_normalized(c.a, wls.v[k])) *
                 c.b;
    return r;
}
// Color * f32 (color.rs:251-270)
RD GpuColor color_scaled(GpuColor c, float s) {
    if (c.kind == PSY_COLOR_XYZ) {
        c.a = c.a * s;
        c.b = c.b * s;
        c.c = c.c * s;
    } else {
        c.b = c.b * s;
    }
    return c;
}
// Color::approximate_energy (color.rs:139-152)
RD float color_energy(const GpuColor& c) {
    if (c.kind == PSY_COLOR_BLACKBODY) {
        float t2 = c.a * c.a;
        return (t2 * t2) * c.b;
    }
    return c.b;
}

// Wyman et al.'s multi-lobe fit of the CIE 1931 observer (color.rs:610-628)
RD float x_1931(float w) {
    float t1 = (w - 442.0f) * (w < 442.0f ? 0.0624f : 0.0374f);
    float t2 = (w - 599.8f) * (w < 599.8f ? 0.0264f : 0.0323f);
    float t3 = (w - 501.1f) * (w < 501.1f ? 0.0490f : 0.0382f);
    return ((0.362f * fast_exp((-0.5f * t1) * t1)) + (1.056f * fast_exp((-0.5f * t2) * t2))) -
           (0.065f * fast_exp((-0.5f * t3) * t3));
}
RD float y_1931(float w) {
    float t1 = (w - 568.8f) * (w < 568.8f ? 0.0213f : 0.0247f);
    float t2 = (w - 530.9f) * (w < 530.9f ? 0.0613f : 0.0322f);
    return (0.821f * fast_exp((-0.5f * t1) * t1)) + (0.286f * fast_exp((-0.5f * t2) * t2));
}
RD float z_1931(float w) {
    float t1 = (w - 437.0f) * (w < 437.0f ? 0.0845f : 0.0278f);
    float t2 = (w - 459.0f) * (w < 459.0f ? 0.0385f : 0.0725f);
    return (1.217f * fast_exp((-0.5f * t1) * t1)) + (0.681f * fast_exp((-0.5f * t2) * t2));
}
// XYZ::from_spectral_sample (color.rs:520-526)
RD V3 xyz_from_spectral(S4 e, float hero) {
    float px[4], py[4], pz[4];
    for (int n = 0; n < 4; n++) {
        float wl = nth_wavelength(hero, n);
        px[n] = x_1931(wl) * e.v[n];
        py[n] = y_1931(wl) * e.v[n];
        pz[n] = z_1931(wl) * e.v[n];
    }
    return v3((((px[0] + px[1]) + px[2]) + px[3]) * 0.75f, (((py[0] + py[1]) + py[2]) + py[3]) * 0.75f,
              (((pz[0] + pz[1]) + pz[2]) + pz[3]) * 0.75f);
}

// ---- lights ----------------------------------------------------------------------

struct LightSample {
    S4 color;
    V3 point, normal;
    float point_err;
    float pdf;
};

RD void rect_corners(const GpuRectLight& L, const Xf& space_inv, V3 p[4]) {
    p[0] = xf_point(space_inv, v3(L.w * 0.5f, L.h * 0.5f, 0.0f));
    p[1] = xf_point(space_inv, v3(L.w * -0.5f, L.h * 0.5f, 0.0f));
    p[2] = xf_point(space_inv, v3(L.w * -0.5f, L.h * -0.5f, 0.0f));
    p[3] = xf_point(space_inv, v3(L.w * 0.5f, L.h * -0.5f, 0.0f));
}

// RectangleLight::sample_pdf (rectangle_light.rs:51-96)
RD float rect_sample_pdf(const GpuRectLight& L, const Xf& space, V3 arr, V3 sample_dir, V3 hit_point) {
    const Xf space_inv = xf_inverse(space);
    V3 p[4];
    rect_corners(L, space_inv, p);
    V3 sp1 = normalize(sub(p[0], arr)), sp2 = normalize(sub(p[1], arr)), sp3 = normalize(sub(p[2], arr)),
       sp4 = normalize(sub(p[3], arr));
    float area_1 = spherical_triangle_solid_angle(sp2, sp1, sp3);
    float area_2 = spherical_triangle_solid_angle(sp4, sp1, sp3);
    V3 normal = xf_normal(space_inv, v3(0.0f, 0.0f, 1.0f));
    if ((area_1 + area_2) < 0.01f) {  // SIMPLE_SAMPLING_THRESHOLD, rectangle_light.rs:21
        float area = triangle_surface_area(p[1], p[0], p[2]) + triangle_surface_area(p[3], p[0], p[2]);
        return (length2(sub(hit_point, arr)) / fabsf(dot(normalize(sample_dir), normalize(normal)))) / area;
    }
    return 1.0f / (area_1 + area_2);
}

// f32::max / f32::min (NaN-ignoring)
RD float fmaxn(float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }
RD float fminn(float a, float b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }

// RectangleLight::sample_from_point (rectangle_light.rs:120-243).  Light
// samples carry colour / area * 0.5 as in the reference.
RD LightSample rect_sample_from_point(const GpuRectLight& L, const Xf& space, V3 arr, float u, float v,
                                      float wavelength) {
    const double surface_area = (double)L.w * (double)L.h;
    const double surface_area_inv = 1.0 / surface_area;
    const Xf space_inv = xf_inverse(space);
    V3 p[4];
    rect_corners(L, space_inv, p);
    V3 sp1 = normalize(sub(p[0], arr)), sp2 = normalize(sub(p[1], arr)), sp3 = normalize(sub(p[2], arr)),
       sp4 = normalize(sub(p[3], arr));
    float area_1 = spherical_triangle_solid_angle(sp2, sp1, sp3);
    float area_2 = spherical_triangle_solid_angle(sp4, sp1, sp3);
    V3 normal = xf_normal(space_inv, v3(0.0f, 0.0f, 1.0f));
    LightSample s;
    s.point_err = 0.0001f;
    s.normal = normal;
    if ((area_1 + area_2) < 0.01f) {
        float sa1 = triangle_surface_area(p[1], p[0], p[2]);
        float sa2 = triangle_surface_area(p[3], p[0], p[2]);
        float thr = sa1 / (sa1 + sa2);
        V3 spt = u < thr ? uniform_sample_triangle(p[1], p[0], p[2], v, u / thr)
                         : uniform_sample_triangle(p[3], p[0], p[2], v, (u - thr) / (1.0f - thr));
        V3 shadow_vec = sub(spt, arr);
        s.color = (color_spectral(L.color, wavelength) * (float)surface_area_inv) * 0.5f;
        s.pdf = (length2(sub(spt, arr)) / fabsf(dot(normalize(shadow_vec), normalize(normal)))) / (sa1 + sa2);
        s.point = spt;
        return s;
    }
    float prob_1 = isinf(area_1) ? 1.0f : (isinf(area_2) ? 0.0f : area_1 / (area_1 + area_2));
    float prob_2 = 1.0f - prob_1;
    V3 shadow_vec = u < prob_1 ? uniform_sample_spherical_triangle(sp2, sp1, sp3, v, u / prob_1)
                               : uniform_sample_spherical_triangle(sp4, sp1, sp3, v, 1.0f - ((u - prob_1) / prob_2));
    V3 arr_local = xf_point(space, arr);
    V3 sv_local = xf_vector(space, shadow_vec);
    sv_local = mul(sv_local, -arr_local.z / sv_local.z);
    V3 spl = add(arr_local, sv_local);
    spl.x = fminn(fmaxn(spl.x, L.w * -0.5f), L.w * 0.5f);
    spl.y = fminn(fmaxn(spl.y, L.h * -0.5f), L.h * 0.5f);
    spl.z = 0.0f;
    s.point = xf_point(space_inv, spl);
    s.pdf = 1.0f / (area_1 + area_2);
    s.color = (color_spectral(L.color, wavelength) * (float)surface_area_inv) * 0.5f;
    return s;
}

// SphereLight::sample_pdf (sphere_light.rs:51-81), arr in world space
RD float sphere_sample_pdf(const GpuSphereLight& L, const Xf& space, V3 arr_world) {
    V3 arr = xf_point(space, arr_world);
    double r = (double)L.radius;
    double d2 = (double)length2(sub(v3(0.0f, 0.0f, 0.0f), arr));
    double d = sqrt(d2);
    if (d > r) {
        double s2 = (r * r) / d2;
        double sin_theta_max2 = s2 < 1.0 ? s2 : 1.0;
        double cos_theta_max = sqrt(1.0 - sin_theta_max2);
        return (float)uniform_sample_cone_pdf(cos_theta_max);
    }
    return (float)(1.0 / (4.0 * PI_D));
}

// SphereLight::sample_from_point (sphere_light.rs:85-192)
RD LightSample sphere_sample_from_point(const GpuSphereLight& L, const Xf& space, V3 arr_world, float u, float v,
                                        float wavelength) {
    V3 arr = xf_point(space, arr_world);
    const Xf inv_space = xf_inverse(space);
    const double r = (double)L.radius;
    const double surface_area_inv = 1.0 / (((4.0 * PI_D) * r) * r);
    V3 zv = sub(v3(0.0f, 0.0f, 0.0f), arr);
    double d2 = (double)length2(zv);
    double d = sqrt(d2);
    V3 z, x, y;
    coordinate_system_from_vector(zv, z, x, y);
    x = normalize(x);
    y = normalize(y);
    z = normalize(z);
    LightSample s;
    s.point_err = len(xf_vector(inv_space, v3((float)r, (float)r, (float)r))) * 0.001f;
    if (d > r) {
        double s2 = (r * r) / d2;
        double sin_theta_max2 = s2 < 1.0 ? s2 : 1.0;
        double cos_theta_max2 = 1.0 - sin_theta_max2;
        double sin_theta_max = sqrt(sin_theta_max2);
        double cos_theta_max = sqrt(cos_theta_max2);
        V3 smp = normalize(uniform_sample_cone(u, v, cos_theta_max));
        double cos_theta = (double)smp.z;
        double cos_theta2 = cos_theta * cos_theta;
        double st2 = 1.0 - cos_theta2;
        double sin_theta2 = st2 > 0.0 ? st2 : 0.0;
        double sin_theta = sqrt(sin_theta2);
        double dd = 1.0 - (((d2 * sin_theta) * sin_theta) / (r * r));
        double cos_a = dd <= 0.0 ? sin_theta_max : (((d / r) * sin_theta2) + (cos_theta * sqrt(dd)));
        double sa2 = 1.0 - (cos_a * cos_a);
        double sin_a = sqrt(sa2 > 0.0 ? sa2 : 0.0);
        double phi = ((double)v * 2.0) * PI_D;
        V3 sv = v3((float)((cos(phi) * sin_a) * r), (float)((sin(phi) * sin_a) * r), (float)(d - (cos_a * r)));
        V3 sample_vec = add(add(mul(x, sv.x), mul(y, sv.y)), mul(z, sv.z));
        V3 normal = normalize(add(arr, sample_vec));
        V3 point = mul(normal, (float)r);
        s.point = xf_point(inv_space, point);
        s.normal = xf_normal(inv_space, normal);
        s.pdf = (float)uniform_sample_cone_pdf(cos_theta_max);
        s.color = color_spectral(L.color, wavelength) * (float)surface_area_inv;
        return s;
    }
    V3 sample_vec = uniform_sample_sphere(u, v);
    V3 normal = normalize(add(arr, sample_vec));
    V3 point = mul(normal, (float)r);
    s.point = xf_point(inv_space, point);
    s.normal = xf_normal(inv_space, normal);
    s.pdf = (float)(1.0 / (4.0 * PI_D));
    s.color = color_spectral(L.color, wavelength) * (float)surface_area_inv;
    return s;
}

// DistantDiskLight::sample_from_point (distant_disk_light.rs:59-85)
RD void distant_sample(const GpuDistantLight& L, float u, float v, float wavelength, S4& col, V3& dir, float& pdf) {
    double r = (double)L.radius;
    double solid_angle_inv = 1.0 / ((2.0 * PI_D) * (1.0 - cos(r)));
    V3 z, x, y;
    coordinate_system_from_vector(neg3(normalize(v3(L.dir[0], L.dir[1], L.dir[2]))), z, x, y);
    double cos_theta_max = cos(r);
    V3 smp = normalize(uniform_sample_cone(u, v, cos_theta_max));
    col = color_spectral(L.color, wavelength) * (float)solid_angle_inv;
    dir = add(add(mul(x, smp.x), mul(y, smp.y)), mul(z, smp.z));
    pdf = (float)uniform_sample_cone_pdf(cos_theta_max);
}

// ---- surface closures (src/shading/surface_closure.rs) ----------------------------

// power_heuristic (src/mis.rs:12-21): returns a / weight.
RD float power_heuristic(float a, float b) {
    if (isinf(a)) return a;
    float a2 = a * a, b2 = b * b;
    float mis_fac = a2 / (a2 + b2);
    return a / mis_fac;
}
RD float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// The normal flipped toward the incoming ray (surface_closure.rs:292-296).
RD void shading_frame(V3 inc, V3 nor, V3 nor_g, V3& nn, V3& fng) {
    if (dot(nor_g, inc) <= 0.0f) {
        nn = normalize(nor);
        fng = nor_g;
    } else {
        nn = neg3(normalize(nor));
        fng = neg3(nor_g);
    }
}
RD float schlick_fresnel_from_fac(float f, float c) {  // :714-718
    float c1 = 1.0f - c;
    float c2 = c1 * c1;
    return f + ((((1.0f - f) * c1) * c2) * c2);
}
RD float ggx_d(float nh, float rough) {  // :611-619
    if (nh <= 0.0f) return 0.0f;
    float rough2 = rough * rough;
    float tmp = 1.0f + ((rough2 - 1.0f) * (nh * nh));
    return rough2 / ((PI_F * tmp) * tmp);
}
RD float ggx_g(float vh, float vn, float rough) {  // :625-631
    if ((vh * vn) <= 0.0f) return 0.0f;
    return 2.0f / (1.0f + sqrtf(1.0f + ((((rough * rough) * (1.0f - (vn * vn))) / (vn * vn)))));
}
RD float half_theta_sample(float u, float rough) {  // :598-606
    float rough2 = rough * rough;
    float top = 1.0f - u;
    float bottom = 1.0f + ((rough2 - 1.0f) * u);
    return sqrtf(top / bottom);
}

// ---- rough dielectric (PSY_SHADER_DIELECTRIC) ------------------------------
// The reference has no transmissive closure; this one follows Walter,
// Marschner, Li and Torrance, "Microfacet Models for Refraction through
// Rough Surfaces" (EGSR 2007): GGX microfacets (the D, G1 and half-vector
// sampling of the GGX closure above), exact Fresnel reflectance choosing
// between reflection and refraction, eqs. 20-21 for the BSDF and 38-41 for
// the sampling densities.  It is kept in throughput form: the refraction
// term carries no (eta_i / eta_o)^2 radiance factor, which cancels between
// entering and leaving for paths that start and end outside.  `ior` is the
// index on the side the geometric normal points away from, `color` tints
// transmitted light, `roughness` is GGX alpha.

// Unpolarised Fresnel reflectance for the cosine on the incident side and
// eta = eta_incident / eta_transmitted; 1 under total internal reflection.
RD float fresnel_dielectric(float cos_i, float eta) {
    const float sin2_t = (eta * eta) * (1.0f - (cos_i * cos_i));
    if (sin2_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(1.0f - sin2_t);
    const float rs = ((eta * cos_i) - cos_t) / ((eta * cos_i) + cos_t);
    const float rp = (cos_i - (eta * cos_t)) / (cos_i + (eta * cos_t));
    return 0.5f * ((rs * rs) + (rp * rp));
}

// f(inc -> out) |cos(out)| and the sampling density of `out`.
RD void dielectric_evaluate(const GpuClosure& c, V3 inc, V3 out, V3 nor, V3 nor_g, float wl, S4& filter,
                            float& pdf) {
    filter = s4_zero();
    pdf = 0.0f;
    const V3 aa = neg3(normalize(inc));
    const V3 bb = normalize(out);
    V3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    const bool entering = dot(nor_g, inc) <= 0.0f;
    const float eta_i = entering ? 1.0f : c.ior;
    const float eta_o = entering ? c.ior : 1.0f;
    const float na = dot(nn, aa);
    const float nb = dot(nn, bb);
    if (na <= 0.0f) return;
    if (dot(fng, bb) >= 0.0f) {
        // Reflection (eq. 20, density eq. 38 times F).
        if (nb <= 0.0f) return;
        const V3 hh = normalize(add(aa, bb));
        const float ha = dot(hh, aa);
        const float hb = dot(hh, bb);
        const float nh = dot(nn, hh);
        if (hb <= 0.0f) return;
        const float f = fresnel_dielectric(clampf(ha, 0.0f, 1.0f), eta_i / eta_o);
        const float d = ggx_d(nh, c.roughness);
        const float g = ggx_g(ha, na, c.roughness) * ggx_g(hb, nb, c.roughness);
        filter = s4_splat(((f * d) * g) / (4.0f * na));
        pdf = ((f * d) * nh) / (4.0f * hb);
        return;
    }
    // Refraction (eq. 21, density eqs. 39-41 times 1 - F).
    if (nb >= 0.0f) return;
    V3 hh = normalize(neg3(add(mul(aa, eta_i), mul(bb, eta_o))));
    if (dot(hh, nn) < 0.0f) hh = neg3(hh);
    const float ha = dot(hh, aa);
    const float hb = dot(hh, bb);
    if (ha <= 0.0f || hb >= 0.0f) return;
    const float nh = dot(nn, hh);
    const float f = fresnel_dielectric(clampf(ha, 0.0f, 1.0f), eta_i / eta_o);
    const float d = ggx_d(nh, c.roughness);
    const float g = ggx_g(ha, na, c.roughness) * ggx_g(hb, nb, c.roughness);
    const float den = (eta_i * ha) + (eta_o * hb);
    const float den2 = den * den;
    const float t = (1.0f - f) * d;
    const float eo2 = eta_o * eta_o;
    filter = color_spectral(c.color, wl) * (((((t * g) * eo2) * ha) * -hb) / (na * den2));
    pdf = (((t * nh) * eo2) * -hb) / den2;
}

// A microfacet normal drawn from D(m) cos(m) as for GGX, then reflection
// with probability F (the third number, w) or refraction.
RD void dielectric_sample(const GpuClosure& c, V3 inc, V3 nor, V3 nor_g, float u, float v, float w, float wl,
                          V3& out, S4& filter, float& pdf) {
    out = v3(0.0f, 0.0f, 0.0f);
    filter = s4_zero();
    pdf = 0.0f;
    V3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    const bool entering = dot(nor_g, inc) <= 0.0f;
    const float eta = entering ? (1.0f / c.ior) : c.ior;
    const float theta_cos = half_theta_sample(u, c.roughness);
    const float theta_sin = sqrtf(1.0f - (theta_cos * theta_cos));
    const float angle = (v * PI_F) * 2.0f;
    const V3 m = normalize(zup_to_vec(v3(psy_cosf(angle) * theta_sin, psy_sinf(angle) * theta_sin, theta_cos), nn));
    const V3 aa = neg3(normalize(inc));
    const float ha = dot(m, aa);
    if (ha <= 0.0f) return;
    const float f = fresnel_dielectric(clampf(ha, 0.0f, 1.0f), eta);
    // A rough surface can send a reflection through the surface or a
    // refraction back out; such samples are rejected, since evaluation
    // attributes every direction to the lobe of its side.
    if (w < f) {
        out = sub(mul(m, 2.0f * ha), aa);
        if (dot(fng, out) < 0.0f) return;
    } else {
        const float sin2_t = (eta * eta) * (1.0f - (ha * ha));
        if (sin2_t >= 1.0f) return;
        const float cos_t = sqrtf(1.0f - sin2_t);
        out = sub(mul(m, (eta * ha) - cos_t), mul(aa, eta));
        if (dot(fng, out) >= 0.0f) return;
    }
    dielectric_evaluate(c, inc, out, nor, nor_g, wl, filter, pdf);
}

// SurfaceClosure::evaluate (Lambert :313-331, GGX :450-521, Emit :657-669)
RD void closure_evaluate(const GpuClosure& c, V3 inc, V3 out, V3 nor, V3 nor_g, float wl, S4& filter, float& pdf) {
    if (c.kind == PSY_SHADER_EMIT) {
        filter = color_spectral(c.color, wl);
        pdf = 1.0f;
        return;
    }
    if (c.kind == PSY_SHADER_LAMBERT) {
        V3 nn, fng;
        shading_frame(inc, nor, nor_g, nn, fng);
        if (dot(fng, out) >= 0.0f) {
            float fac = fmaxn(dot(nn, normalize(out)), 0.0f) * INV_PI_F;
            filter = color_spectral(c.color, wl) * fac;
            pdf = fac;
        } else {
            filter = s4_zero();
            pdf = 0.0f;
        }
        return;
    }
    if (c.kind == PSY_SHADER_DIELECTRIC) {
        dielectric_evaluate(c, inc, out, nor, nor_g, wl, filter, pdf);
        return;
    }
    V3 aa = neg3(normalize(inc));
    V3 bb = normalize(out);
    V3 hh = normalize(add(aa, bb));
    V3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    if (dot(nn, aa) < 0.0f || dot(nn, bb) < 0.0f || dot(fng, bb) < 0.0f) {
        filter = s4_zero();
        pdf = 0.0f;
        return;
    }
    float na = clampf(dot(nn, aa), -1.0f, 1.0f);
    float nb = clampf(dot(nn, bb), -1.0f, 1.0f);
    float ha = clampf(dot(hh, aa), -1.0f, 1.0f);
    float hb = clampf(dot(hh, bb), -1.0f, 1.0f);
    float nh = clampf(dot(nn, hh), -1.0f, 1.0f);
    S4 ss = color_spectral(c.color, wl);
    float rev = 1.0f - c.fresnel;
    S4 col_f;
    for (int k = 0; k < 4; k++) col_f.v[k] = lerpf(schlick_fresnel_from_fac(ss.v[k], hb), ss.v[k], rev);
    if (c.roughness == 0.0f) {
        filter = col_f;
        pdf = 0.0f;
        return;
    }
    float dist = ggx_d(nh, c.roughness) / na;
    float g1 = ggx_g(ha, na, c.roughness);
    float g2 = ggx_g(hb, nb, c.roughness);
    filter = (col_f * ((dist * g1) * g2)) * INV_PI_F;
    pdf = dist * INV_PI_F;
}

// SurfaceClosure::sample (Lambert :284-311, GGX :418-448, Emit :641-655);
// `w`, a third sample number, is read by the dielectric only.
RD void closure_sample(const GpuClosure& c, V3 inc, V3 nor, V3 nor_g, float u, float v, float w, float wl, V3& out,
                       S4& filter, float& pdf) {
    if (c.kind == PSY_SHADER_EMIT) {
        out = v3(0.0f, 0.0f, 0.0f);
        filter = color_spectral(c.color, wl);
        pdf = 1.0f;
        return;
    }
    if (c.kind == PSY_SHADER_DIELECTRIC) {
        dielectric_sample(c, inc, nor, nor_g, u, v, w, wl, out, filter, pdf);
        return;
    }
    V3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    if (c.kind == PSY_SHADER_LAMBERT) {
        V3 dir = cosine_sample_hemisphere(u, v);
        float p = dir.z * INV_PI_F;
        out = zup_to_vec(dir, nn);
        if (dot(fng, out) >= 0.0f) {
            filter = color_spectral(c.color, wl) * p;
            pdf = p;
        } else {
            filter = s4_zero();
            pdf = 0.0f;
        }
        return;
    }
    float theta_cos = half_theta_sample(u, c.roughness);
    float theta_sin = sqrtf(1.0f - (theta_cos * theta_cos));
    float angle = (v * PI_F) * 2.0f;
    V3 half_dir = v3(psy_cosf(angle) * theta_sin, psy_sinf(angle) * theta_sin, theta_cos);
    half_dir = normalize(zup_to_vec(half_dir, nn));
    out = sub(inc, mul(mul(half_dir, 2.0f), dot(inc, half_dir)));
    if (dot(fng, out) >= 0.0f) {
        closure_evaluate(c, inc, out, nor, nor_g, wl, filter, pdf);
    } else {
        filter = s4_zero();
        pdf = 0.0f;
    }
}

// sphere_lambert (surface_closure.rs:348-373)
RD float sphere_lambert(float nlcos, float rcos) {
    float nlsin = sqrtf(1.0f - (nlcos * nlcos));
    float rsin2 = 1.0f - (rcos * rcos);
    float rsin = sqrtf(rsin2);
    float ysin = rcos / nlsin;
    float ycos2 = 1.0f - (ysin * ysin);
    float ycos = sqrtf(ycos2);
    float g = ((((-2.0f * nlsin) * rcos) * ycos) + H_PI_F - psy_asinf(ysin)) + (ysin * ycos);
    float h = nlcos * ((ycos * sqrtf(rsin2 - ycos2)) + (rsin2 * psy_asinf(ycos / rsin)));
    float nl = psy_acosf(nlcos);
    float r = psy_acosf(rcos);
    if (nl < (H_PI_F - r)) return nlcos * rsin2;
    if (nl < H_PI_F) return ((nlcos * rsin2) + g) - h;
    if (nl < (H_PI_F + r)) return (g + h) * INV_PI_F;
    return 0.0f;
}

// SurfaceClosure::estimate_eval_over_sphere_light (Lambert :333-407, GGX
// :523-592); the reference leaves it unimplemented for Emit and has no
// dielectric.
RD float closure_estimate_sphere_light(const GpuClosure& c, V3 inc, V3 to_light_center, float r2, V3 nor, V3 nor_g) {
    if (c.kind == PSY_SHADER_LAMBERT) {
        float dist2 = length2(to_light_center);
        if (dist2 <= r2) return fminn(r2 / dist2, 4.0f);
        float sin_theta_max2 = fminn(r2 / dist2, 1.0f);
        float cos_theta_max = sqrtf(1.0f - sin_theta_max2);
        V3 vv = normalize(to_light_center);
        V3 nn = dot(nor_g, inc) <= 0.0f ? normalize(nor) : neg3(normalize(nor));
        float cos_nv = fminn(fmaxn(dot(nn, vv), -1.0f), 1.0f);
        return sphere_lambert(cos_nv, cos_theta_max);
    }
    if (c.kind == PSY_SHADER_GGX) {
        float dist2 = length2(to_light_center);
        float sin_theta_max2 = fminn(r2 / dist2, 1.0f);
        float cos_theta_max = sqrtf(1.0f - sin_theta_max2);
        V3 nn = dot(nor, inc) < 0.0f ? normalize(nor) : neg3(normalize(nor));
        V3 aa = neg3(normalize(inc));
        V3 bb = normalize(to_light_center);
        float theta = psy_acosf(cos_theta_max);
        V3 hh = normalize(add(aa, bb));
        float nh = clampf(dot(nn, hh), -1.0f, 1.0f);
        float fac = ggx_d(nh, fminn(1.0f, sqrtf(c.roughness) + ((2.0f * theta) / PI_F)));
        return (fac * fminn(1.0f, 1.0f - cos_theta_max)) * INV_PI_F;
    }
    if (c.kind == PSY_SHADER_DIELECTRIC) {
        // Lights on either side count: the light's solid angle over 2 pi.
        float dist2 = length2(to_light_center);
        if (dist2 <= r2) return 1.0f;
        float sin_theta_max2 = fminn(r2 / dist2, 1.0f);
        return 1.0f - sqrtf(1.0f - sin_theta_max2);
    }
    return 0.0f;
}

}  // namespace rd
}  // namespace psy
