// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Surface closures: restates src/shading/surface_closure.rs (Lambert, GGX,
// Emit: sample, evaluate, estimate_eval_over_sphere_light), this build's
// rough dielectric (no reference counterpart), and
// src/shading/mod.rs (SimpleSurfaceShader), and the MIS weight of
// src/mis.rs.
#pragma once

#include "psyref_color.h"
#include "psyref_math.h"
#include "psyref_sample.h"

namespace psyref {

static const float INV_PI = 1.0f / PI_32;
static const float H_PI = PI_32 / 2.0f;

// CLOSURE_DIELECTRIC has no counterpart in the reference (its closures are
// Lambert, GGX and Emit, surface_closure.rs:19-30); it is this build's rough
// dielectric, restated here from its specification below.
enum ClosureKind : uint32_t { CLOSURE_LAMBERT = 0, CLOSURE_GGX = 1, CLOSURE_EMIT = 2, CLOSURE_DIELECTRIC = 3 };

struct Closure {
    uint32_t kind = CLOSURE_LAMBERT;
    Color color;
    float roughness = 0, fresnel = 0;
    float ior = 0;  // dielectric only
};

// mis.rs:12-21
static inline float power_heuristic(float a, float b) {
    if (std::isinf(a)) return a;
    float a2 = a * a, b2 = b * b;
    float mis_fac = a2 / (a2 + b2);
    return a / mis_fac;
}

// f32::clamp(-1, 1)
static inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
// f32::max / min (NaN-ignoring)
static inline float fmaxn(float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }
static inline float fminn(float a, float b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }

// The shading frame shared by every closure: the normal flipped toward the
// incoming ray (surface_closure.rs:292-296 and repeats).
static inline void shading_frame(Vec3 inc, Vec3 nor, Vec3 nor_g, Vec3& nn, Vec3& flipped_nor_g) {
    if (dot(nor_g, inc) <= 0.0f) {
        nn = normalized(nor);
        flipped_nor_g = nor_g;
    } else {
        nn = neg(normalized(nor));
        flipped_nor_g = neg(nor_g);
    }
}

// surface_closure.rs:714-718
static inline float schlick_fresnel_from_fac(float f, float c) {
    float c1 = 1.0f - c;
    float c2 = c1 * c1;
    return f + ((((1.0f - f) * c1) * c2) * c2);
}
static inline float ggx_d(float nh, float rough) {  // :611-619
    if (nh <= 0.0f) return 0.0f;
    float rough2 = rough * rough;
    float tmp = 1.0f + ((rough2 - 1.0f) * (nh * nh));
    return rough2 / ((PI_32 * tmp) * tmp);
}
static inline float ggx_g(float vh, float vn, float rough) {  // :625-631
    if ((vh * vn) <= 0.0f) return 0.0f;
    return 2.0f / (1.0f + std::sqrt(1.0f + ((((rough * rough) * (1.0f - (vn * vn))) / (vn * vn)))));
}
static inline float half_theta_sample(float u, float rough) {  // :598-606
    float rough2 = rough * rough;
    float top = 1.0f - u;
    float bottom = 1.0f + ((rough2 - 1.0f) * u);
    return std::sqrt(top / bottom);
}

// ---- rough dielectric (no reference counterpart) ---------------------------
// Walter, Marschner, Li and Torrance, "Microfacet Models for Refraction
// through Rough Surfaces" (EGSR 2007): GGX microfacets with the GGX
// closure's D, G1 and half-vector sampling, exact unpolarised Fresnel
// reflectance choosing reflection or refraction, the BSDF of eqs. 20-21 and
// the densities of eqs. 38-41, in throughput form (no (eta_i/eta_o)^2
// radiance factor).  `ior` is the index on the side the geometric normal
// points away from; `color` tints refracted light.
static inline float fresnel_dielectric(float cos_i, float eta) {
    const float sin2_t = (eta * eta) * (1.0f - (cos_i * cos_i));
    if (sin2_t >= 1.0f) return 1.0f;  // total internal reflection
    const float cos_t = std::sqrt(1.0f - sin2_t);
    const float rs = ((eta * cos_i) - cos_t) / ((eta * cos_i) + cos_t);
    const float rp = (cos_i - (eta * cos_t)) / (cos_i + (eta * cos_t));
    return 0.5f * ((rs * rs) + (rp * rp));
}

static inline void dielectric_evaluate(const Closure& c, Vec3 inc, Vec3 out, Vec3 nor, Vec3 nor_g, float wl,
                                       Vec4& filter, float& pdf) {
    filter = Vec4();
    pdf = 0.0f;
    const Vec3 aa = neg(normalized(inc));
    const Vec3 bb = normalized(out);
    Vec3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    const bool entering = dot(nor_g, inc) <= 0.0f;
    const float eta_i = entering ? 1.0f : c.ior;
    const float eta_o = entering ? c.ior : 1.0f;
    const float na = dot(nn, aa);
    const float nb = dot(nn, bb);
    if (na <= 0.0f) return;
    if (dot(fng, bb) >= 0.0f) {
        // reflection: f = F D G / (4 |i.n| |o.n|), density F D (m.n) / (4 |o.m|)
        if (nb <= 0.0f) return;
        const Vec3 hh = normalized(aa + bb);
        const float ha = dot(hh, aa);
        const float hb = dot(hh, bb);
        const float nh = dot(nn, hh);
        if (hb <= 0.0f) return;
        const float f = fresnel_dielectric(clampf(ha, 0.0f, 1.0f), eta_i / eta_o);
        const float d = ggx_d(nh, c.roughness);
        const float g = ggx_g(ha, na, c.roughness) * ggx_g(hb, nb, c.roughness);
        const float v = ((f * d) * g) / (4.0f * na);
        filter = Vec4(v, v, v, v);
        pdf = ((f * d) * nh) / (4.0f * hb);
        return;
    }
    // refraction: f = |i.m||o.m| eta_o^2 (1-F) D G / (|i.n||o.n| (eta_i i.m + eta_o o.m)^2),
    // density (1-F) D (m.n) eta_o^2 |o.m| / (eta_i i.m + eta_o o.m)^2
    if (nb >= 0.0f) return;
    Vec3 hh = normalized(neg((aa * eta_i) + (bb * eta_o)));
    if (dot(hh, nn) < 0.0f) hh = neg(hh);
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
    filter = c.color.to_spectral_sample(wl) * (((((t * g) * eo2) * ha) * -hb) / (na * den2));
    pdf = (((t * nh) * eo2) * -hb) / den2;
}

static inline void dielectric_sample(const Closure& c, Vec3 inc, Vec3 nor, Vec3 nor_g, float u, float v, float w,
                                     float wl, Vec3& out, Vec4& filter, float& pdf) {
    out = Vec3(0, 0, 0);
    filter = Vec4();
    pdf = 0.0f;
    Vec3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    const bool entering = dot(nor_g, inc) <= 0.0f;
    const float eta = entering ? (1.0f / c.ior) : c.ior;
    const float theta_cos = half_theta_sample(u, c.roughness);
    const float theta_sin = std::sqrt(1.0f - (theta_cos * theta_cos));
    const float angle = (v * PI_32) * 2.0f;
    const Vec3 m = normalized(zup_to_vec(Vec3(std::cos(angle) * theta_sin, std::sin(angle) * theta_sin, theta_cos), nn));
    const Vec3 aa = neg(normalized(inc));
    const float ha = dot(m, aa);
    if (ha <= 0.0f) return;
    const float f = fresnel_dielectric(clampf(ha, 0.0f, 1.0f), eta);
    // reflections through the surface and refractions back out are
    // rejected: evaluation attributes each direction to its side's lobe
    if (w < f) {
        out = (m * (2.0f * ha)) - aa;
        if (dot(fng, out) < 0.0f) return;
    } else {
        const float sin2_t = (eta * eta) * (1.0f - (ha * ha));
        if (sin2_t >= 1.0f) return;
        const float cos_t = std::sqrt(1.0f - sin2_t);
        out = (m * ((eta * ha) - cos_t)) - (aa * eta);
        if (dot(fng, out) >= 0.0f) return;
    }
    dielectric_evaluate(c, inc, out, nor, nor_g, wl, filter, pdf);
}

// ---- evaluate -------------------------------------------------------------
static inline void closure_evaluate(const Closure& c, Vec3 inc, Vec3 out, Vec3 nor, Vec3 nor_g, float wl, Vec4& filter,
                                    float& pdf) {
    if (c.kind == CLOSURE_EMIT) {  // :657-669
        filter = c.color.to_spectral_sample(wl);
        pdf = 1.0f;
        return;
    }
    if (c.kind == CLOSURE_LAMBERT) {  // :313-331
        Vec3 nn, fng;
        shading_frame(inc, nor, nor_g, nn, fng);
        if (dot(fng, out) >= 0.0f) {
            float fac = fmaxn(dot(nn, normalized(out)), 0.0f) * INV_PI;
            filter = c.color.to_spectral_sample(wl) * fac;
            pdf = fac;
        } else {
            filter = Vec4();
            pdf = 0.0f;
        }
        return;
    }
    // GGX (:450-521)
    if (c.kind == CLOSURE_DIELECTRIC) {
        dielectric_evaluate(c, inc, out, nor, nor_g, wl, filter, pdf);
        return;
    }
    Vec3 aa = neg(normalized(inc));
    Vec3 bb = normalized(out);
    Vec3 hh = normalized(aa + bb);
    Vec3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    if (dot(nn, aa) < 0.0f || dot(nn, bb) < 0.0f || dot(fng, bb) < 0.0f) {
        filter = Vec4();
        pdf = 0.0f;
        return;
    }
    float na = clampf(dot(nn, aa), -1.0f, 1.0f);
    float nb = clampf(dot(nn, bb), -1.0f, 1.0f);
    float ha = clampf(dot(hh, aa), -1.0f, 1.0f);
    float hb = clampf(dot(hh, bb), -1.0f, 1.0f);
    float nh = clampf(dot(nn, hh), -1.0f, 1.0f);
    Vec4 ss = c.color.to_spectral_sample(wl);
    float rev = 1.0f - c.fresnel;
    Vec4 col_f;
    for (int k = 0; k < 4; k++) col_f[k] = lerp(schlick_fresnel_from_fac(ss[k], hb), ss[k], rev);
    if (c.roughness == 0.0f) {
        filter = col_f;
        pdf = 0.0f;
        return;
    }
    float dist = ggx_d(nh, c.roughness) / na;
    float g1 = ggx_g(ha, na, c.roughness);
    float g2 = ggx_g(hb, nb, c.roughness);
    filter = (col_f * ((dist * g1) * g2)) * INV_PI;
    pdf = dist * INV_PI;
}

// ---- sample ----------------------------------------------------------------
// `w`, a third sample number, is read by the dielectric only.
static inline void closure_sample(const Closure& c, Vec3 inc, Vec3 nor, Vec3 nor_g, float u, float v, float w,
                                  float wl, Vec3& out, Vec4& filter, float& pdf) {
    if (c.kind == CLOSURE_EMIT) {  // :641-655
        out = Vec3(0, 0, 0);
        filter = c.color.to_spectral_sample(wl);
        pdf = 1.0f;
        return;
    }
    if (c.kind == CLOSURE_DIELECTRIC) {
        dielectric_sample(c, inc, nor, nor_g, u, v, w, wl, out, filter, pdf);
        return;
    }
    Vec3 nn, fng;
    shading_frame(inc, nor, nor_g, nn, fng);
    if (c.kind == CLOSURE_LAMBERT) {  // :284-311
        Vec3 dir = cosine_sample_hemisphere(u, v);
        float p = dir.z * INV_PI;
        out = zup_to_vec(dir, nn);
        if (dot(fng, out) >= 0.0f) {
            filter = c.color.to_spectral_sample(wl) * p;
            pdf = p;
        } else {
            filter = Vec4();
            pdf = 0.0f;
        }
        return;
    }
    // GGX (:418-448)
    float theta_cos = half_theta_sample(u, c.roughness);
    float theta_sin = std::sqrt(1.0f - (theta_cos * theta_cos));
    float angle = (v * PI_32) * 2.0f;
    Vec3 half_dir(std::cos(angle) * theta_sin, std::sin(angle) * theta_sin, theta_cos);
    half_dir = normalized(zup_to_vec(half_dir, nn));
    out = inc - ((half_dir * 2.0f) * dot(inc, half_dir));
    if (dot(fng, out) >= 0.0f) {
        closure_evaluate(c, inc, out, nor, nor_g, wl, filter, pdf);
    } else {
        filter = Vec4();
        pdf = 0.0f;
    }
}

// ---- estimate_eval_over_sphere_light ----------------------------------------
static inline float sphere_lambert(float nlcos, float rcos) {  // :348-373
    float nlsin = std::sqrt(1.0f - (nlcos * nlcos));
    float rsin2 = 1.0f - (rcos * rcos);
    float rsin = std::sqrt(rsin2);
    float ysin = rcos / nlsin;
    float ycos2 = 1.0f - (ysin * ysin);
    float ycos = std::sqrt(ycos2);
    float g = ((((-2.0f * nlsin) * rcos) * ycos) + H_PI - std::asin(ysin)) + (ysin * ycos);
    float h = nlcos * ((ycos * std::sqrt(rsin2 - ycos2)) + (rsin2 * std::asin(ycos / rsin)));
    float nl = std::acos(nlcos);
    float r = std::acos(rcos);
    if (nl < (H_PI - r)) return nlcos * rsin2;
    if (nl < H_PI) return ((nlcos * rsin2) + g) - h;
    if (nl < (H_PI + r)) return (g + h) * INV_PI;
    return 0.0f;
}

static inline float closure_estimate_sphere_light(const Closure& c, Vec3 inc, Vec3 to_light_center, float r2,
                                                  Vec3 nor, Vec3 nor_g) {
    if (c.kind == CLOSURE_LAMBERT) {  // :333-407
        float dist2 = length2(to_light_center);
        if (dist2 <= r2) return fminn(r2 / dist2, 4.0f);
        float sin_theta_max2 = fminn(r2 / dist2, 1.0f);
        float cos_theta_max = std::sqrt(1.0f - sin_theta_max2);
        Vec3 vv = normalized(to_light_center);
        Vec3 nn = dot(nor_g, inc) <= 0.0f ? normalized(nor) : neg(normalized(nor));
        float cos_nv = fminn(fmaxn(dot(nn, vv), -1.0f), 1.0f);
        return sphere_lambert(cos_nv, cos_theta_max);
    }
    if (c.kind == CLOSURE_GGX) {  // :523-592
        float dist2 = length2(to_light_center);
        float sin_theta_max2 = fminn(r2 / dist2, 1.0f);
        float cos_theta_max = std::sqrt(1.0f - sin_theta_max2);
        Vec3 nn = dot(nor, inc) < 0.0f ? normalized(nor) : neg(normalized(nor));
        Vec3 aa = neg(normalized(inc));
        Vec3 bb = normalized(to_light_center);
        float theta = std::acos(cos_theta_max);
        Vec3 hh = normalized(aa + bb);
        float nh = clampf(dot(nn, hh), -1.0f, 1.0f);
        float fac = ggx_d(nh, fminn(1.0f, std::sqrt(c.roughness) + ((2.0f * theta) / PI_32)));
        return (fac * fminn(1.0f, 1.0f - cos_theta_max)) * INV_PI;
    }
    if (c.kind == CLOSURE_DIELECTRIC) {
        // lights on either side: the light's solid angle over 2 pi
        float dist2 = length2(to_light_center);
        if (dist2 <= r2) return 1.0f;
        float sin_theta_max2 = fminn(r2 / dist2, 1.0f);
        return 1.0f - std::sqrt(1.0f - sin_theta_max2);
    }
    return 0.0f;  // Emit: unimplemented!() in the reference; never reached (emitters end paths)
}

}  // namespace psyref
