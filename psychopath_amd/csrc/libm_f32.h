// Single-precision sine, cosine, arcsine and arccosine with the rounding of
// the host C library the reference links (Rust's f32::sin/cos/asin/acos call
// the platform libm).  asinf/acosf follow Sun fdlibm's float algorithms
// (e_asinf.c with Moshier's Cephes polynomial, e_acosf.c) and agree bit for
// bit with the host's on every input in [-1, 1]; sinf/cosf follow the
// double-precision polynomial method of Arm's optimized-routines (Nagy,
// 2018) and agree bit for bit with the host's on every input with
// |x| <= 8, which covers every angle the renderer takes (beyond 120 they
// fall back to a double-precision evaluation).  The device library's own
// versions differ in the last bit on many inputs: the light tree's closure
// estimates cancel enough digits to turn that into different light
// choices, and glossy reflection between convex surfaces amplifies it into
// different paths.  Compile without FMA contraction.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "psymath.h"

namespace psy {

PSY_HD uint32_t lm_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
PSY_HD float lm_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

PSY_HD float psy_asinf(float x) {
    const float pio2_hi = 1.57079637050628662109375f;
    const float pio2_lo = -4.37113900018624283e-8f;
    const float pio4_hi = 0.785398185253143310546875f;
    const float p0 = 1.666675248e-1f, p1 = 7.495297643e-2f, p2 = 4.547037598e-2f, p3 = 2.417951451e-2f,
                p4 = 4.216630880e-2f;
    const int32_t hx = (int32_t)lm_bits(x);
    const int32_t ix = hx & 0x7fffffff;
    if (ix == 0x3f800000) return x * pio2_hi + x * pio2_lo;
    if (ix > 0x3f800000) return (x - x) / (x - x);
    if (ix < 0x3f000000) {
        if (ix < 0x32000000) return x;
        const float t = x * x;
        const float w = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
        return x + x * w;
    }
    float w = 1.0f - fabsf(x);
    float t = w * 0.5f;
    float p = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
    const float s = sqrtf(t);
    if (ix >= 0x3F79999A) {
        t = pio2_hi - (2.0f * (s + s * p) - pio2_lo);
    } else {
        w = lm_float(lm_bits(s) & 0xfffff000u);
        const float c = (t - w * w) / (s + w);
        const float r = p;
        p = 2.0f * s * r - (pio2_lo - 2.0f * c);
        const float q = pio4_hi - 2.0f * w;
        t = pio4_hi - (p - q);
    }
    return hx > 0 ? t : -t;
}

PSY_HD float psy_acosf(float x) {
    const float pi = 3.1415925026e+00f, pio2_hi = 1.5707962513e+00f, pio2_lo = 7.5497894159e-08f;
    const float pS0 = 1.6666667163e-01f, pS1 = -3.2556581497e-01f, pS2 = 2.0121252537e-01f,
                pS3 = -4.0055535734e-02f, pS4 = 7.9153501429e-04f, pS5 = 3.4793309169e-05f;
    const float qS1 = -2.4033949375e+00f, qS2 = 2.0209457874e+00f, qS3 = -6.8828397989e-01f,
                qS4 = 7.7038154006e-02f;
    const int32_t hx = (int32_t)lm_bits(x);
    const int32_t ix = hx & 0x7fffffff;
    if (ix == 0x3f800000) return hx > 0 ? 0.0f : pi + 2.0f * pio2_lo;
    if (ix > 0x3f800000) return (x - x) / (x - x);
    if (ix < 0x3f000000) {
        if (ix <= 0x32800000) return pio2_hi + pio2_lo;
        const float z = x * x;
        const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        const float q = 1.0f + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        const float r = p / q;
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if (hx < 0) {
        const float z = (1.0f + x) * 0.5f;
        const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        const float q = 1.0f + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        const float s = sqrtf(z);
        const float r = p / q;
        const float w = r * s - pio2_lo;
        return pi - 2.0f * (s + w);
    }
    const float z = (1.0f - x) * 0.5f;
    const float s = sqrtf(z);
    const float df = lm_float(lm_bits(s) & 0xfffff000u);
    const float c = (z - df * df) / (s + df);
    const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    const float q = 1.0f + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    const float r = p / q;
    const float w = r * s + c;
    return 2.0f * (df + w);
}

// ---- sinf / cosf ------------------------------------------------------------

// Coefficients of the method: minimax fits of cos (C0..C4) and sin
// (S1..S3) on [-pi/4, pi/4], evaluated in double precision.  Quadrants where
// cos changes sign use the negated cos coefficients.
constexpr double LM_HPI_INV = 0x1.45F306DC9C883p+23;  // 2/pi, scaled by 2^24
constexpr double LM_HPI = 0x1.921FB54442D18p0;        // pi/2
constexpr double LM_C0 = 0x1p0, LM_C1 = -0x1.ffffffd0c621cp-2, LM_C2 = 0x1.55553e1068f19p-5,
                 LM_C3 = -0x1.6c087e89a359dp-10, LM_C4 = 0x1.99343027bf8c3p-16;
constexpr double LM_S1 = -0x1.555545995a603p-3, LM_S2 = 0x1.1107605230bc4p-7, LM_S3 = -0x1.994eb3774cf24p-13;

PSY_HD uint32_t lm_abstop12(float x) { return (lm_bits(x) >> 20) & 0x7ff; }

// sin (n even) or cos (n odd) of x in [-pi/4, pi/4]; x2 = x * x.
PSY_HD float lm_sincos_poly(double x, double x2, bool neg, int n) {
    if ((n & 1) == 0) {
        const double x3 = x * x2;
        const double s1 = LM_S2 + x2 * LM_S3;
        const double x7 = x3 * x2;
        const double s = x + x3 * LM_S1;
        return (float)(s + x7 * s1);
    }
    const double c0 = neg ? -LM_C0 : LM_C0, c1 = neg ? -LM_C1 : LM_C1, c2 = neg ? -LM_C2 : LM_C2,
                 c3 = neg ? -LM_C3 : LM_C3, c4 = neg ? -LM_C4 : LM_C4;
    const double x4 = x2 * x2;
    const double p2 = c3 + x2 * c4;
    const double p1 = c0 + x2 * c1;
    const double x6 = x4 * x2;
    const double c = p1 + x4 * c2;
    return (float)(c + x6 * p2);
}

// x - n * pi/2 with n = round(x * 2/pi), for |x| < 120.
PSY_HD double lm_reduce(double x, int& n) {
    const double r = x * LM_HPI_INV;
    n = ((int32_t)r + 0x800000) >> 24;
    return x - n * LM_HPI;
}
PSY_HD double lm_quadrant_sign(int n) { return ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0; }

PSY_HD float psy_sinf(float y) {
    const uint32_t top = lm_abstop12(y);
    if (top < lm_abstop12(0x1.921FB6p-1f)) {
        const double x = y;
        if (top < lm_abstop12(0x1p-12f)) return y;
        return lm_sincos_poly(x, x * x, false, 0);
    }
    if (top < lm_abstop12(120.0f)) {
        int n;
        const double x = lm_reduce((double)y, n);
        return lm_sincos_poly(x * lm_quadrant_sign(n), x * x, (n & 2) != 0, n);
    }
    return (float)sin((double)y);
}

PSY_HD float psy_cosf(float y) {
    const uint32_t top = lm_abstop12(y);
    if (top < lm_abstop12(0x1.921FB6p-1f)) {
        const double x = y;
        if (top < lm_abstop12(0x1p-12f)) return 1.0f;
        return lm_sincos_poly(x, x * x, false, 1);
    }
    if (top < lm_abstop12(120.0f)) {
        int n;
        const double x = lm_reduce((double)y, n);
        return lm_sincos_poly(x * lm_quadrant_sign(n), x * x, (n & 2) != 0, n ^ 1);
    }
    return (float)cos((double)y);
}

}  // namespace psy
