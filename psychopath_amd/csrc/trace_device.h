// Device-side geometry shared by the traversal kernels: the BVH4 box test,
// the watertight triangle test and transform helpers.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/psy_rt.h"
#include "layout.h"
#include "light_samples.h"
#include "psymath.h"

namespace psy {

// The 4-wide slab test of bbox4.rs:62-97 with minps/maxps semantics.
__device__ __forceinline__ bool slab(float lox, float hix, float loy, float hiy, float loz, float hiz, V3 o, V3 di,
                                     float tmax) {
    float t1x = (lox - o.x) * di.x;
    float t1y = (loy - o.y) * di.y;
    float t1z = (loz - o.z) * di.z;
    float t2x = (hix - o.x) * di.x;
    float t2y = (hiy - o.y) * di.y;
    float t2z = (hiz - o.z) * di.z;
    float tfx = fmax_sse(t1x, t2x), tfy = fmax_sse(t1y, t2y), tfz = fmax_sse(t1z, t2z);
    float tnx = fmin_sse(t1x, t2x), tny = fmin_sse(t1y, t2y), tnz = fmin_sse(t1z, t2z);
    float far_t = fmin_sse(fmin_sse(tfx, fmin_sse(tfy, tfz)) * BBOX_MAXT_ADJUST, tmax);
    float near_t = fmax_sse(fmax_sse(tnx, tny), fmax_sse(tnz, 0.0f));
    return near_t < far_t;
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float alpha) {
    float ia = 1.0f - alpha;
    return make_float4((a.x * ia) + (b.x * alpha), (a.y * ia) + (b.y * alpha), (a.z * ia) + (b.z * alpha),
                       (a.w * ia) + (b.w * alpha));
}

__device__ __forceinline__ uint32_t octant_of(V3 di) {
    return (uint32_t)(di.x < 0.0f) | ((uint32_t)(di.y < 0.0f) << 1) | ((uint32_t)(di.z < 0.0f) << 2);
}

__device__ __forceinline__ float max_abs3(float a, float b, float c) {
    a = fabsf(a); b = fabsf(b); c = fabsf(c);
    if (a > b && a > c) return a;
    if (b > c) return b;
    return c;
}

struct TriPre {
    int kx, ky, kz;
    float sx, sy, sz;
};

// RayTriPrecompute::new (triangle.rs:15-44)
__device__ __forceinline__ TriPre tri_precompute(V3 d) {
    TriPre p;
    float xa = fabsf(d.x), ya = fabsf(d.y), za = fabsf(d.z);
    if (xa > ya && xa > za) { p.kx = 1; p.ky = 2; p.kz = 0; }
    else if (ya > za) { p.kx = 2; p.ky = 0; p.kz = 1; }
    else { p.kx = 0; p.ky = 1; p.kz = 2; }
    float dx = comp(d, p.kx), dy = comp(d, p.ky), dz = comp(d, p.kz);
    p.sx = dx / dz;
    p.sy = dy / dz;
    p.sz = 1.0f / dz;
    return p;
}

// triangle::intersect_ray (triangle.rs:58-149)
__device__ __forceinline__ bool tri_test(V3 o, const TriPre& pre, float tmax, V3 v0, V3 v1, V3 v2, float& t_out,
                                         float& b0o, float& b1o, float& b2o) {
    V3 p0 = sub(v0, o), p1 = sub(v1, o), p2 = sub(v2, o);
    float p0z_ = comp(p0, pre.kz), p1z_ = comp(p1, pre.kz), p2z_ = comp(p2, pre.kz);
    float p0x = comp(p0, pre.kx) - (pre.sx * p0z_);
    float p0y = comp(p0, pre.ky) - (pre.sy * p0z_);
    float p1x = comp(p1, pre.kx) - (pre.sx * p1z_);
    float p1y = comp(p1, pre.ky) - (pre.sy * p1z_);
    float p2x = comp(p2, pre.kx) - (pre.sx * p2z_);
    float p2y = comp(p2, pre.ky) - (pre.sy * p2z_);

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

    float p0z = pre.sz * p0z_;
    float p1z = pre.sz * p1z_;
    float p2z = pre.sz * p2z_;
    float t_scaled = ((e0 * p0z) + (e1 * p1z)) + (e2 * p2z);
    if ((det > 0.0f && (t_scaled <= 0.0f || t_scaled > (tmax * det))) ||
        (det < 0.0f && (t_scaled >= 0.0f || t_scaled < (tmax * det))))
        return false;

    float inv_det = 1.0f / det;
    float b0 = e0 * inv_det, b1 = e1 * inv_det, b2 = e2 * inv_det;
    float t = t_scaled * inv_det;

    float max_zt = max_abs3(p0z, p1z, p2z);
    float dz = fp_gamma(3) * max_zt;
    float max_xt = max_abs3(p0x, p1x, p2x);
    float max_yt = max_abs3(p0y, p1y, p2y);
    float dx = fp_gamma(5) * (max_xt + max_zt);
    float dy = fp_gamma(5) * (max_yt + max_zt);
    float de = 2.0f * (((fp_gamma(2) * max_xt) * max_yt) + ((dy * max_xt) + (dx * max_yt)));
    float max_e = max_abs3(e0, e1, e2);
    float dt = (3.0f * (((fp_gamma(3) * max_e) * max_zt) + ((de * max_zt) + (dz * max_e)))) * fabsf(inv_det);
    if (t <= dt) return false;
    t_out = t; b0o = b0; b1o = b1; b2o = b2;
    return true;
}

// The instance's world-to-object transform at `time` (lerp_slice over its
// time samples, src/lerp.rs:20-35); xf_count must be at least 1.
__device__ __forceinline__ Xf load_xf(const GpuXform* x);
__device__ __forceinline__ Xf instance_xf(const GpuXform* xforms, const GpuInstance& inst, float time) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(inst.xf_count, time, i1, a2)) return load_xf(xforms + inst.xf_start + i1);
    return xf_lerp(load_xf(xforms + inst.xf_start + i1), load_xf(xforms + inst.xf_start + i1 + 1), a2);
}

__device__ __forceinline__ Xf load_xf(const GpuXform* x) {
    const float4* p = reinterpret_cast<const float4*>(x);
    float4 a = p[0], b = p[1], c = p[2];
    Xf r;
    r.m[0][0] = a.x; r.m[0][1] = a.y; r.m[0][2] = a.z; r.m[0][3] = a.w;
    r.m[1][0] = b.x; r.m[1][1] = b.y; r.m[1][2] = b.z; r.m[1][3] = b.w;
    r.m[2][0] = c.x; r.m[2][1] = c.y; r.m[2][2] = c.z; r.m[2][3] = c.w;
    return r;
}


// Light surfaces as traversal leaves.  Rectangle (rectangle_light.rs:257-
// 338): the two world-space triangles (p1,p2,p3), (p3,p4,p1) of the corners
// mapped through the inverse of the lerped transform, the first hit with
// t < tmax wins; prim is the triangle.  Sphere (sphere_light.rs:206-282): the
// quadratic in object space with the object-space origin `lo` and direction
// `ld`.  Returns true on a hit.
__device__ __forceinline__ bool light_leaf_test(const GpuRectLight* rects, const GpuSphereLight* spheres,
                                                const float* light_samples, float time, const GpuInstance& inst,
                                                const Xf& x, V3 o, const TriPre& pre, V3 lo, V3 ld, float tmax,
                                                float& t_out, float& b0, float& b1, float& b2, uint32_t& prim) {
    if (inst.kind == PSY_OBJECT_RECT_LIGHT) {
        const GpuRectLight L = rect_at(rects[inst.mesh], light_samples, time);
        const Xf inv = xf_inverse(x);
        const float hw = L.w * 0.5f, hh = L.h * 0.5f, nw = L.w * -0.5f, nh = L.h * -0.5f;
        const V3 p1 = xf_point(inv, v3(hw, hh, 0.0f));
        const V3 p2 = xf_point(inv, v3(nw, hh, 0.0f));
        const V3 p3 = xf_point(inv, v3(nw, nh, 0.0f));
        const V3 p4 = xf_point(inv, v3(hw, nh, 0.0f));
        float t, c0, c1, c2;
        if (tri_test(o, pre, tmax, p1, p2, p3, t, c0, c1, c2) && t < tmax) {
            t_out = t; b0 = c0; b1 = c1; b2 = c2; prim = 0;
            return true;
        }
        if (tri_test(o, pre, tmax, p3, p4, p1, t, c0, c1, c2) && t < tmax) {
            t_out = t; b0 = c0; b1 = c1; b2 = c2; prim = 1;
            return true;
        }
        return false;
    }
    const float r = sphere_at(spheres[inst.mesh], light_samples, time).radius;
    const float qa = dot(ld, ld);
    const float qb = 2.0f * dot(ld, lo);
    const float qc = dot(lo, lo) - (r * r);
    float disc = (qb * qb) - ((4.0f * qa) * qc);
    if (disc < 0.0f) return false;
    disc = sqrtf(disc);
    const float q = qb < 0.0f ? -0.5f * (qb - disc) : -0.5f * (qb + disc);
    float t0 = q / qa;
    float t1 = q != 0.0f ? qc / q : tmax;
    if (t0 > t1) {
        float tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (t0 > tmax || t1 <= 0.0f) return false;
    float t;
    if (t0 > 0.0f) t = t0;
    else if (t1 <= tmax) t = t1;
    else return false;
    t_out = t; b0 = 0.0f; b1 = 0.0f; b2 = 0.0f; prim = 0;
    return true;
}

}  // namespace psy
