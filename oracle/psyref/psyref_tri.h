// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Watertight ray-triangle test: restates src/surface/triangle.rs.
#pragma once

#include "psyref_math.h"

namespace psyref {

// triangle.rs:9-45
struct RayTriPrecompute {
    int kx, ky, kz;
    float sx, sy, sz;
    explicit RayTriPrecompute(Vec3 d) {
        float xabs = std::fabs(d.x), yabs = std::fabs(d.y), zabs = std::fabs(d.z);
        if (xabs > yabs && xabs > zabs) { kx = 1; ky = 2; kz = 0; }
        else if (yabs > zabs) { kx = 2; ky = 0; kz = 1; }
        else { kx = 0; ky = 1; kz = 2; }
        float dx = d[kx], dy = d[ky], dz = d[kz];
        sx = dx / dz;
        sy = dy / dz;
        sz = 1.0f / dz;
    }
};

// triangle.rs:171-183
static inline float max_abs_3(float a, float b, float c) {
    a = std::fabs(a); b = std::fabs(b); c = std::fabs(c);
    if (a > b && a > c) return a;
    if (b > c) return b;
    return c;
}

struct TriHit { float t, b0, b1, b2; };

// triangle.rs:47-149
static inline bool intersect_triangle(Vec3 orig, const RayTriPrecompute& pre, float max_t, Vec3 t0, Vec3 t1, Vec3 t2,
                                      TriHit& out) {
    Vec3 p0 = t0 - orig, p1 = t1 - orig, p2 = t2 - orig;
    float p0x = p0[pre.kx] - (pre.sx * p0[pre.kz]);
    float p0y = p0[pre.ky] - (pre.sy * p0[pre.kz]);
    float p1x = p1[pre.kx] - (pre.sx * p1[pre.kz]);
    float p1y = p1[pre.ky] - (pre.sy * p1[pre.kz]);
    float p2x = p2[pre.kx] - (pre.sx * p2[pre.kz]);
    float p2y = p2[pre.ky] - (pre.sy * p2[pre.kz]);

    float e0 = (p1x * p2y) - (p1y * p2x);
    float e1 = (p2x * p0y) - (p2y * p0x);
    float e2 = (p0x * p1y) - (p0y * p1x);
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        e0 = (float)(((double)p1x * (double)p2y) - ((double)p1y * (double)p2x));
        e1 = (float)(((double)p2x * (double)p0y) - ((double)p2y * (double)p0x));
        e2 = (float)(((double)p0x * (double)p1y) - ((double)p0y * (double)p1x));
    }
    if ((e0 < 0.0f || e1 < 0.0f || e2 < 0.0f) && (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f)) return false;

    float det = (e0 + e1) + e2;
    if (det == 0.0f) return false;

    float p0z = pre.sz * p0[pre.kz];
    float p1z = pre.sz * p1[pre.kz];
    float p2z = pre.sz * p2[pre.kz];
    float t_scaled = ((e0 * p0z) + (e1 * p1z)) + (e2 * p2z);

    if ((det > 0.0f && (t_scaled <= 0.0f || t_scaled > (max_t * det))) ||
        (det < 0.0f && (t_scaled >= 0.0f || t_scaled < (max_t * det))))
        return false;

    float inv_det = 1.0f / det;
    float b0 = e0 * inv_det, b1 = e1 * inv_det, b2 = e2 * inv_det;
    float t = t_scaled * inv_det;

    float max_zt = max_abs_3(p0z, p1z, p2z);
    float dz = fp_gamma(3) * max_zt;
    float max_xt = max_abs_3(p0x, p1x, p2x);
    float max_yt = max_abs_3(p0y, p1y, p2y);
    float dx = fp_gamma(5) * (max_xt + max_zt);
    float dy = fp_gamma(5) * (max_yt + max_zt);
    float de = 2.0f * (((fp_gamma(2) * max_xt) * max_yt) + ((dy * max_xt) + (dx * max_yt)));
    float max_e = max_abs_3(e0, e1, e2);
    float dt = (3.0f * (((fp_gamma(3) * max_e) * max_zt) + ((de * max_zt) + (dz * max_e)))) * std::fabs(inv_det);
    if (t <= dt) return false;

    out.t = t; out.b0 = b0; out.b1 = b1; out.b2 = b2;
    return true;
}

// triangle.rs:151-169
static inline void surface_point(Vec3 t0, Vec3 t1, Vec3 t2, float b0, float b1, float b2, Vec3& pos, float& pos_err) {
    pos = ((t0 * b0) + (t1 * b1)) + (t2 * b2);
    pos_err = max_element((((vabs(t0) * b0) + (vabs(t1) * b1)) + (vabs(t2) * b2)) * fp_gamma(7));
}

}  // namespace psyref
