// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Sampling: the Faure-permuted Halton sequence of sub_crates/halton/build.rs,
// the renderer's sample streams (src/renderer.rs:698-737) with Halton in
// place of sobol_burley, and the Monte Carlo mappings of
// src/sampling/monte_carlo.rs.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "psyref_math.h"

namespace psyref {

// ---- sub_crates/halton/build.rs -------------------------------------------
struct Halton {
    static const int NUM_DIMENSIONS = 128;
    struct Dim {
        uint32_t base, pow_base, max_power;
        std::vector<uint16_t> perm;
    };
    std::vector<Dim> dims;

    Halton() {
        std::vector<uint32_t> primes;
        for (uint32_t c = 2; primes.size() < (size_t)NUM_DIMENSIONS; c++) {
            bool p = true;
            for (uint32_t i = 2; i < c; i++)
                if (c % i == 0) { p = false; break; }
            if (p) primes.push_back(c);
        }
        // Faure permutations for bases 0..=largest prime (build.rs:246-280).
        std::vector<std::vector<uint32_t>> faure;
        for (uint32_t b = 0; b <= primes.back(); b++) {
            std::vector<uint32_t> perm;
            if (b < 2) perm = {0};
            else if (b == 2) perm = {0, 1};
            else if (b & 1) {
                uint32_t c = (b - 1) / 2;
                for (uint32_t i = 0; i < b; i++) {
                    if (i == c) { perm.push_back(c); continue; }
                    uint32_t f = faure[b - 1][i - (i > c ? 1 : 0)];
                    perm.push_back(f + (f >= c ? 1 : 0));
                }
            } else {
                uint32_t c = b / 2;
                for (uint32_t i = 0; i < b; i++) perm.push_back(i < c ? 2 * faure[c][i] : 2 * faure[c][i - c] + 1);
            }
            faure.push_back(perm);
        }
        dims.resize(NUM_DIMENSIONS);
        for (int i = 1; i < NUM_DIMENSIONS; i++) {
            Dim& d = dims[i];
            uint64_t base = primes[i];
            uint64_t digits = 1, pow_base = base;
            while (pow_base * base <= 500) { pow_base *= base; digits++; }
            uint64_t max_power = pow_base;
            while (max_power * pow_base < (1ull << 32)) max_power *= pow_base;
            d.base = (uint32_t)base;
            d.pow_base = (uint32_t)pow_base;
            d.max_power = (uint32_t)max_power;
            for (uint64_t j = 0; j < pow_base; j++) {  // invert (build.rs:282-291)
                uint64_t index = j, result = 0;
                for (uint64_t k = 0; k < digits; k++) {
                    uint64_t rem = index % base;
                    index /= base;
                    result = result * base + faure[base][rem];
                }
                d.perm.push_back((uint16_t)result);
            }
        }
    }

    // The generated `sample(index, dimension)` (build.rs:101-228).
    float sample(uint32_t dimension, uint32_t index) const {
        if (dimension == 0) {
            index = (index << 16) | (index >> 16);
            index = ((index & 0x00ff00ffu) << 8) | ((index & 0xff00ff00u) >> 8);
            index = ((index & 0x0f0f0f0fu) << 4) | ((index & 0xf0f0f0f0u) >> 4);
            index = ((index & 0x33333333u) << 2) | ((index & 0xccccccccu) >> 2);
            index = ((index & 0x55555555u) << 1) | ((index & 0xaaaaaaaau) >> 1);
            return (float)index * (1.0f / 4294967296.0f);
        }
        const Dim& d = dims[dimension];
        uint32_t power = d.max_power / d.pow_base;
        uint32_t sum = (uint32_t)d.perm[index % d.pow_base] * power;
        uint32_t div = 1;
        while (power / d.pow_base > 1) {
            div *= d.pow_base;
            power /= d.pow_base;
            sum += (uint32_t)d.perm[(index / div) % d.pow_base] * power;
        }
        sum += (uint32_t)d.perm[(index / (div * d.pow_base)) % d.pow_base];
        return (float)sum * (0.999999940395355224609375f / (float)d.max_power);
    }
};

const Halton& halton();

// get_sample_4d (renderer.rs:698-725) with the Halton sequence in place of
// sobol_burley: dimension sets below 32 use Halton dimensions 4*set..4*set+3
// at index `i`, each shifted by a per-pixel, per-dimension offset
// hash_u32_to_f32(dim, seed) modulo 1 (a Cranley-Patterson rotation, which
// decorrelates pixels as the Owen scrambling of the Sobol sampler did);
// higher sets use the reference's hash fallback unchanged.
static inline void get_sample_4d(uint32_t i, uint32_t dimension_set, uint32_t px, uint32_t py, uint32_t seed,
                                 float out[4]) {
    uint32_t s = px ^ (py << 16) ^ (seed * 0x736caf6fu);
    if (dimension_set < (uint32_t)Halton::NUM_DIMENSIONS / 4) {
        for (uint32_t k = 0; k < 4; k++) {
            uint32_t dim = dimension_set * 4 + k;
            float v = halton().sample(dim, i) + hash_u32_to_f32(dim, s);
            if (v >= 1.0f) v = v - 1.0f;
            out[k] = v;
        }
    } else {
        for (uint32_t k = 0; k < 4; k++) out[k] = hash_u32_to_f32((dimension_set * 4 + k) ^ (i << 16), s);
    }
}

// golden_ratio_sample (renderer.rs:728-737)
static inline float golden_ratio_sample(uint32_t i, uint32_t scramble) {
    uint32_t n = (i + hash_u32(scramble, 0)) * 2654435769u;
    return (float)n * (1.0f / 4294967296.0f);
}

// ---- src/sampling/monte_carlo.rs -------------------------------------------
static const float QPI_32 = 0.785398163397448309615660845819875721f;
static const float PI_32 = 3.14159265358979323846264338327950288f;
static const double PI_64 = 3.14159265358979323846264338327950288;

// square_to_circle (monte_carlo.rs:10-32)
static inline void square_to_circle(float x, float y, float& ox, float& oy) {
    if (x == 0.0f && y == 0.0f) { ox = 0.0f; oy = 0.0f; return; }
    float radius, angle;
    if (x > std::fabs(y)) { radius = x; angle = QPI_32 * (y / x); }
    else if (y > std::fabs(x)) { radius = y; angle = QPI_32 * (2.0f - (x / y)); }
    else if (x < -std::fabs(y)) { radius = -x; angle = QPI_32 * (4.0f + (y / x)); }
    else { radius = -y; angle = QPI_32 * (6.0f - (x / y)); }
    ox = radius * std::cos(angle);
    oy = radius * std::sin(angle);
}
// cosine_sample_hemisphere (monte_carlo.rs:34-38)
static inline Vec3 cosine_sample_hemisphere(float u, float v) {
    float uu, vv;
    square_to_circle((u * 2.0f) - 1.0f, (v * 2.0f) - 1.0f, uu, vv);
    float t = 1.0f - ((uu * uu) + (vv * vv));
    float z = std::sqrt(t > 0.0f ? t : 0.0f);
    return Vec3(uu, vv, z);
}
// uniform_sample_sphere (monte_carlo.rs:49-56)
static inline Vec3 uniform_sample_sphere(float u, float v) {
    float z = 1.0f - (2.0f * u);
    float t = 1.0f - (z * z);
    float r = std::sqrt(t > 0.0f ? t : 0.0f);
    float phi = (2.0f * PI_32) * v;
    return Vec3(r * std::cos(phi), r * std::sin(phi), z);
}
// uniform_sample_cone (monte_carlo.rs:64-73)
static inline Vec3 uniform_sample_cone(float u, float v, double cos_theta_max) {
    double cos_theta = (1.0 - (double)u) + ((double)u * cos_theta_max);
    double sin_theta = std::sqrt(1.0 - (cos_theta * cos_theta));
    double phi = ((double)v * 2.0) * PI_64;
    return Vec3((float)(std::cos(phi) * sin_theta), (float)(std::sin(phi) * sin_theta), (float)cos_theta);
}
static inline double uniform_sample_cone_pdf(double cos_theta_max) {
    return 1.0 / ((2.0 * PI_64) * (1.0 - cos_theta_max));
}
// uniform_sample_triangle (monte_carlo.rs:82-89)
static inline Vec3 uniform_sample_triangle(Vec3 va, Vec3 vb, Vec3 vc, float i, float j) {
    float isqrt = std::sqrt(i);
    float a = 1.0f - isqrt;
    float b = isqrt * (1.0f - j);
    float c = j * isqrt;
    return ((va * a) + (vb * b)) + (vc * c);
}
// triangle_surface_area (monte_carlo.rs:92-94)
static inline float triangle_surface_area(Vec3 p0, Vec3 p1, Vec3 p2) {
    return 0.5f * length(cross(p1 - p0, p2 - p0));
}
// f64::max(-1.0).min(1.0): NaN -> -1.
static inline double clamp11(double x) { return (x != x) ? -1.0 : (x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x)); }
static inline float clamp11f(float x) {
    // f32::max(-1.0).min(1.0) (IEEE maxNum/minNum: NaN -> -1 -> -1)
    float m = (x != x) ? -1.0f : (x > -1.0f ? x : -1.0f);
    return m < 1.0f ? m : 1.0f;
}
// spherical_triangle_solid_angle (monte_carlo.rs:99-131)
static inline float spherical_triangle_solid_angle(Vec3 va, Vec3 vb, Vec3 vc) {
    double cos_a = (double)clamp11f(dot(vb, vc));
    double cos_b = (double)clamp11f(dot(vc, va));
    double cos_c = (double)clamp11f(dot(va, vb));
    double sin_a = std::sqrt(1.0 - (cos_a * cos_a));
    double sin_b = std::sqrt(1.0 - (cos_b * cos_b));
    double sin_c = std::sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return 0.0f;
    double cos_va = clamp11((cos_a - (cos_b * cos_c)) / (sin_b * sin_c));
    double cos_vb = clamp11((cos_b - (cos_c * cos_a)) / (sin_c * sin_a));
    double cos_vc = clamp11((cos_c - (cos_a * cos_b)) / (sin_a * sin_b));
    return (float)(((std::acos(cos_va) + std::acos(cos_vb)) + std::acos(cos_vc)) - PI_64);
}
// uniform_sample_spherical_triangle (monte_carlo.rs:135-200), Arvo's method.
static inline Vec3 uniform_sample_spherical_triangle(Vec3 va, Vec3 vb, Vec3 vc, float i, float j) {
    double cos_a = (double)clamp11f(dot(vb, vc));
    double cos_b = (double)clamp11f(dot(vc, va));
    double cos_c = (double)clamp11f(dot(va, vb));
    double sin_a = std::sqrt(1.0 - (cos_a * cos_a));
    double sin_b = std::sqrt(1.0 - (cos_b * cos_b));
    double sin_c = std::sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return va;
    double cos_va = clamp11((cos_a - (cos_b * cos_c)) / (sin_b * sin_c));
    double cos_vb = clamp11((cos_b - (cos_c * cos_a)) / (sin_c * sin_a));
    double cos_vc = clamp11((cos_c - (cos_a * cos_b)) / (sin_a * sin_b));
    double sin_va = std::sqrt(1.0 - (cos_va * cos_va));
    double ang_va = std::acos(cos_va), ang_vb = std::acos(cos_vb), ang_vc = std::acos(cos_vc);
    double area = ((ang_va + ang_vb) + ang_vc) - PI_64;
    double area_2 = area * (double)i;
    double s = std::sin(area_2 - ang_va);
    double t = std::cos(area_2 - ang_va);
    double u = t - cos_va;
    double v = s + (sin_va * cos_c);
    double q_top = (((v * t) - (u * s)) * cos_va) - v;
    double q_bottom = ((v * s) + (u * t)) * sin_va;
    double q = q_top / q_bottom;
    Vec3 vc_2 = (va * (float)q) + (normalized(vc - (va * dot(vc, va))) * (float)std::sqrt(1.0 - (q * q)));
    float z = 1.0f - (j * (1.0f - dot(vc_2, vb)));
    return (vb * z) + (normalized(vc_2 - (vb * dot(vc_2, vb))) * std::sqrt(1.0f - (z * z)));
}

// ---- src/math.rs:19-54 ------------------------------------------------------
// f32::signum: 1.0 for +0 and positives, -1.0 for -0 and negatives.
static inline float signum_f32(float x) { return std::signbit(x) ? -1.0f : 1.0f; }
static inline void coordinate_system_from_vector(Vec3 v, Vec3& v1, Vec3& v2, Vec3& v3) {
    float sign = signum_f32(v.z);
    float a = -1.0f / (sign + v.z);
    float b = (v.x * v.y) * a;
    v1 = v;
    v2 = Vec3(1.0f + (((sign * v.x) * v.x) * a), sign * b, -sign * v.x);
    v3 = Vec3(b, sign + ((v.y * v.y) * a), -v.y);
}
static inline Vec3 zup_to_vec(Vec3 from, Vec3 toz) {
    Vec3 tz, tx, ty;
    coordinate_system_from_vector(normalized(toz), tz, tx, ty);
    return ((tx * from.x) + (ty * from.y)) + (tz * from.z);
}

}  // namespace psyref
