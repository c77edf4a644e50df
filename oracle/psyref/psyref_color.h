// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Spectral colour: restates src/color.rs (hero wavelengths, Color,
// to_spectral_sample, Planck's law, XYZ::from_spectral_sample with the Wyman
// CIE fits), sub_crates/spectral_upsampling/src/meng.rs
// (spectrum_xyz_to_p_4) and the fastapprox 0.3 `fast::exp` the colour code
// calls (src/math.rs:146-148; Mineiro's fastpow2).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifndef PSY_MENG_STORAGE
#define PSY_MENG_STORAGE static
#endif
#include "../../psychopath_amd/csrc/meng_tables.h"

namespace psyref {

struct Vec4 {
    float v[4];
    Vec4() : v{0, 0, 0, 0} {}
    Vec4(float a, float b, float c, float d) : v{a, b, c, d} {}
    static Vec4 splat(float a) { return Vec4(a, a, a, a); }
    float operator[](int i) const { return v[i]; }
    float& operator[](int i) { return v[i]; }
};
static inline Vec4 operator+(Vec4 a, Vec4 b) { return Vec4(a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]); }
static inline Vec4 operator-(Vec4 a, Vec4 b) { return Vec4(a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3]); }
static inline Vec4 operator*(Vec4 a, Vec4 b) { return Vec4(a[0] * b[0], a[1] * b[1], a[2] * b[2], a[3] * b[3]); }
static inline Vec4 operator*(Vec4 a, float s) { return Vec4(a[0] * s, a[1] * s, a[2] * s, a[3] * s); }
static inline Vec4 operator/(Vec4 a, float s) { return Vec4(a[0] / s, a[1] / s, a[2] / s, a[3] / s); }
static inline Vec4 operator/(Vec4 a, Vec4 b) { return Vec4(a[0] / b[0], a[1] / b[1], a[2] / b[2], a[3] / b[3]); }
// glam Vec4::max_element on SSE2: max(max(x,z), max(y,w)).
static inline float max_element4(Vec4 a) {
    float m0 = a[0] > a[2] ? a[0] : a[2];
    float m1 = a[1] > a[3] ? a[1] : a[3];
    return m0 > m1 ? m0 : m1;
}

// ---- fastapprox 0.3 (Mineiro) -------------------------------------------
static inline uint32_t f32_to_u32_saturating(float f) {
    if (!(f > 0.0f)) return 0;  // NaN and negatives saturate to 0 (Rust `as u32`)
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
static inline int32_t f32_to_i32_saturating(float f) {
    if (f != f) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483648.0f) return INT32_MAX;
    return (int32_t)f;
}
static inline float fast_pow2(float p) {
    float offset = p < 0.0f ? 1.0f : 0.0f;
    float clipp = p < -126.0f ? -126.0f : p;
    int32_t w = f32_to_i32_saturating(clipp);
    float z = (clipp - (float)w) + offset;
    float v = (float)(1 << 23) * (((clipp + 121.2740575f) + (27.7280233f / (4.84252568f - z))) - (1.49012907f * z));
    uint32_t bits = f32_to_u32_saturating(v);
    float out;
    std::memcpy(&out, &bits, 4);
    return out;
}
static inline float fast_exp(float p) { return fast_pow2(1.442695040f * p); }

// ---- color.rs:15-43 ------------------------------------------------------
static const float WL_MIN = 380.0f, WL_MAX = 700.0f;
static const float WL_RANGE = WL_MAX - WL_MIN;
static const float WL_RANGE_Q = WL_RANGE / 4.0f;
static inline float map_0_1_to_wavelength(float n) { return (n * WL_RANGE) + WL_MIN; }
static inline float nth_wavelength(float hero, int n) {
    float wl = hero + (WL_RANGE_Q * (float)n);
    return wl > WL_MAX ? wl - WL_RANGE : wl;
}
static inline Vec4 wavelengths(float hero) {
    return Vec4(nth_wavelength(hero, 0), nth_wavelength(hero, 1), nth_wavelength(hero, 2), nth_wavelength(hero, 3));
}

// ---- meng.rs:177-328 -----------------------------------------------------
static inline Vec4 spectrum_xyz_to_p_4(Vec4 lambdas, float x, float y, float z) {
    using namespace psy_meng;
    float inv_norm = (x + y) + z;
    float norm = 1.0f / inv_norm;
    if (!(norm < FLT_MAX)) return Vec4();
    float xx = x * norm, yy = y * norm;
    float u = ((MAT_XY_TO_UV[0] * xx) + (MAT_XY_TO_UV[1] * yy)) + MAT_XY_TO_UV[2];
    float v = ((MAT_XY_TO_UV[3] * xx) + (MAT_XY_TO_UV[4] * yy)) + MAT_XY_TO_UV[5];
    if (u < 0.0f || u >= (float)GRID_WIDTH || v < 0.0f || v >= (float)GRID_HEIGHT) return Vec4();
    int ui = (int)u, vi = (int)v;
    const int* cell = GRID[ui + GRID_WIDTH * vi];
    bool inside = cell[0] != 0;
    int num = cell[1];
    const int* idx = cell + 2;
    if (num == 0) return Vec4();

    Vec4 sb = ((lambdas - Vec4::splat(SAMPLE_MIN)) / (SAMPLE_MAX - SAMPLE_MIN)) * ((float)NUM_SAMPLES - 1.0f);
    int sb0[4], sb1[4];
    for (int k = 0; k < 4; k++) {
        sb0[k] = (int)sb[k];
        sb1[k] = sb0[k] + 1 < NUM_SAMPLES - 1 ? sb0[k] + 1 : NUM_SAMPLES - 1;
    }
    Vec4 sbf = sb - Vec4((float)sb0[0], (float)sb0[1], (float)sb0[2], (float)sb0[3]);
    Vec4 p[6];
    for (int i = 0; i < num; i++) {
        const float* spectrum = POINT_SPECTRUM[idx[i]];
        Vec4 p0(spectrum[sb0[0]], spectrum[sb0[1]], spectrum[sb0[2]], spectrum[sb0[3]]);
        Vec4 p1(spectrum[sb1[0]], spectrum[sb1[1]], spectrum[sb1[2]], spectrum[sb1[3]]);
        p[i] = (p0 * (Vec4::splat(1.0f) - sbf)) + (p1 * sbf);
    }

    Vec4 interp;
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
        for (int i = 0; i < num - 1; i++) {
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
            interp = ((p[0] * wf) + (p[i + 1] * vf)) + (p[i == num - 2 ? 1 : i + 2] * uf);
            break;
        }
    }
    return interp * inv_norm;
}

// color.rs:603-606
static inline Vec4 xyz_to_spectrum_4(float x, float y, float z, Vec4 wls) {
    return spectrum_xyz_to_p_4(wls, x, y, z) * Vec4::splat(1.0f / psy_meng::EQUAL_ENERGY_REFLECTANCE);
}

// color.rs:336-383
static inline float plancks_law(float temperature, float wavelength) {
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
    return e > 0.0f ? e : 0.0f;  // f32::max(0.0): NaN -> 0
}
static inline float plancks_law_normalized(float temperature, float wavelength) {
    float t2 = temperature * temperature;
    return (plancks_law(temperature, wavelength) * 4.0e7f) / (t2 * t2);
}

// color.rs:46-137
enum ColorKind : uint32_t { COLOR_XYZ = 0, COLOR_BLACKBODY = 1, COLOR_TEMPERATURE = 2 };
struct Color {
    uint32_t kind = COLOR_XYZ;
    float a = 0, b = 0, c = 0;  // XYZ: x, y, z; Blackbody/Temperature: temperature, factor
    Vec4 to_spectral_sample(float hero) const {
        Vec4 wls = wavelengths(hero);
        switch (kind) {
            case COLOR_XYZ: return xyz_to_spectrum_4(a, b, c, wls);
            case COLOR_BLACKBODY:
                return Vec4(plancks_law(a, wls[0]) * b, plancks_law(a, wls[1]) * b, plancks_law(a, wls[2]) * b,
                            plancks_law(a, wls[3]) * b);
            default:
                return Vec4(plancks_law_normalized(a, wls[0]) * b, plancks_law_normalized(a, wls[1]) * b,
                            plancks_law_normalized(a, wls[2]) * b, plancks_law_normalized(a, wls[3]) * b);
        }
    }
    float approximate_energy() const {
        switch (kind) {
            case COLOR_XYZ: return b;
            case COLOR_BLACKBODY: {
                float t2 = a * a;
                return (t2 * t2) * b;
            }
            default: return b;
        }
    }
    Color scaled(float s) const {  // Color * f32 (color.rs:251-270)
        Color r = *this;
        if (kind == COLOR_XYZ) { r.a = a * s; r.b = b * s; r.c = c * s; }
        else r.b = b * s;
        return r;
    }
};

// Lerp for Color (color.rs:284-333): component-wise within one
// representation; the reader refuses lists that mix representations.
static inline Color lerp(const Color& x, const Color& y, float alpha) {
    const float inv = 1.0f - alpha;
    Color r = x;
    r.a = (x.a * inv) + (y.a * alpha);
    r.b = (x.b * inv) + (y.b * alpha);
    if (x.kind == COLOR_XYZ) r.c = (x.c * inv) + (y.c * alpha);
    return r;
}

// Light::approximate_energy over a colour list: the mean of the colours'
// estimates, summed from 0 in list order (rectangle_light.rs:249-254).
static inline float mean_energy(const std::vector<Color>& cs) {
    float e = 0.0f;
    for (const Color& c : cs) e = e + c.approximate_energy();
    return e / (float)cs.size();
}

// color.rs:610-628
static inline float x_1931(float w) {
    float t1 = (w - 442.0f) * (w < 442.0f ? 0.0624f : 0.0374f);
    float t2 = (w - 599.8f) * (w < 599.8f ? 0.0264f : 0.0323f);
    float t3 = (w - 501.1f) * (w < 501.1f ? 0.0490f : 0.0382f);
    return ((0.362f * fast_exp((-0.5f * t1) * t1)) + (1.056f * fast_exp((-0.5f * t2) * t2))) -
           (0.065f * fast_exp((-0.5f * t3) * t3));
}
static inline float y_1931(float w) {
    float t1 = (w - 568.8f) * (w < 568.8f ? 0.0213f : 0.0247f);
    float t2 = (w - 530.9f) * (w < 530.9f ? 0.0613f : 0.0322f);
    return (0.821f * fast_exp((-0.5f * t1) * t1)) + (0.286f * fast_exp((-0.5f * t2) * t2));
}
static inline float z_1931(float w) {
    float t1 = (w - 437.0f) * (w < 437.0f ? 0.0845f : 0.0278f);
    float t2 = (w - 459.0f) * (w < 459.0f ? 0.0385f : 0.0725f);
    return (1.217f * fast_exp((-0.5f * t1) * t1)) + (0.681f * fast_exp((-0.5f * t2) * t2));
}

struct XYZ {
    float x = 0, y = 0, z = 0;
};
// XYZ::from_spectral_sample (color.rs:520-526)
static inline XYZ xyz_from_spectral_sample(Vec4 e, float hero) {
    XYZ acc;
    XYZ parts[4];
    for (int n = 0; n < 4; n++) {
        float wl = nth_wavelength(hero, n);
        parts[n].x = x_1931(wl) * e[n];
        parts[n].y = y_1931(wl) * e[n];
        parts[n].z = z_1931(wl) * e[n];
    }
    acc.x = (((parts[0].x + parts[1].x) + parts[2].x) + parts[3].x) * 0.75f;
    acc.y = (((parts[0].y + parts[1].y) + parts[2].y) + parts[3].y) * 0.75f;
    acc.z = (((parts[0].z + parts[1].z) + parts[2].z) + parts[3].z) * 0.75f;
    return acc;
}

}  // namespace psyref
