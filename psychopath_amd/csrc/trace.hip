// BVH4 ray traversal for gfx950.
//
// One lane per ray.  Each lane walks the instance BVH4 and, at an instance
// leaf, the mesh BVH4 in that instance's space: the same depth-first order
// the reference's ray-stream traversal gives every ray
// (src/accel/bvh4.rs:101-184, src/tracer.rs:87-150), the same box test
// (src/bbox4.rs:63-97), motion lerps (src/lerp.rs:20-35) and watertight
// triangle test in world space (src/surface/triangle.rs:58-149,
// src/surface/triangle_mesh.rs:174-260).  The pending-node stack lives in LDS
// (one column per lane) with a private overflow area.
#include <hip/hip_runtime.h>

#include "layout.h"
#include "psymath.h"
#include "trace.h"

namespace psy {

namespace {

constexpr int BLOCK = TRACE_BLOCK;
constexpr int LDS_DEPTH = TRACE_LDS_DEPTH;
constexpr int SPILL_DEPTH = TRACE_STACK_CAPACITY - TRACE_LDS_DEPTH;

// Pending-node stack: entries [0, LDS_DEPTH) in this lane's LDS column,
// the rest in a private overflow array.  The depth counter stays in a
// register.
#define STACK_PUSH(v)                                                  \
    do {                                                               \
        if (sp < LDS_DEPTH) lds_col[sp * BLOCK] = (v);                 \
        else spill[sp - LDS_DEPTH] = (v);                              \
        sp++;                                                          \
    } while (0)
#define STACK_POP() (--sp, sp < LDS_DEPTH ? lds_col[sp * BLOCK] : spill[sp - LDS_DEPTH])

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

__device__ __forceinline__ Xf load_xf(const GpuXform* x) {
    const float4* p = reinterpret_cast<const float4*>(x);
    float4 a = p[0], b = p[1], c = p[2];
    Xf r;
    r.m[0][0] = a.x; r.m[0][1] = a.y; r.m[0][2] = a.z; r.m[0][3] = a.w;
    r.m[1][0] = b.x; r.m[1][1] = b.y; r.m[1][2] = b.z; r.m[1][3] = b.w;
    r.m[2][0] = c.x; r.m[2][1] = c.y; r.m[2][2] = c.z; r.m[2][3] = c.w;
    return r;
}

template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void trace_kernel(TraceArgs a) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * BLOCK];
    const uint64_t ray = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (ray >= a.n) return;

    const V3 o = v3(a.ox[ray], a.oy[ray], a.oz[ray]);
    const V3 d = v3(a.dx[ray], a.dy[ray], a.dz[ray]);
    const float time = a.time[ray];
    float tmax = a.max_t[ray];
    const bool occl = a.flags != nullptr && (a.flags[ray] & 1u) != 0;

    // Top-level local ray: update_local with the identity (tracer.rs:65-70).
    const Xf ident = xf_identity();
    const V3 top_lo = xf_point(ident, o);
    const V3 top_ld = xf_vector(ident, d);
    const V3 top_di = v3(1.0f / top_ld.x, 1.0f / top_ld.y, 1.0f / top_ld.z);
    const uint32_t top_oct = octant_of(top_di);
    const TriPre pre = tri_precompute(d);

    V3 lo = top_lo, di = top_di;
    uint32_t oct = top_oct;

    uint32_t* const lds_col = lds_stack + threadIdx.x;
    uint32_t spill[SPILL_DEPTH];
    int sp = 0;

    // Instance state.
    bool in_inst = false;
    int inst_base = 0;
    bool has_space = false;
    Xf space;
    uint32_t inst_id = PSY_NONE;
    uint64_t tri_base = 0;
    uint32_t tri_stride = 0, tri_samples = 1;

    // Result.
    uint8_t kind = PSY_HIT_MISS;
    float ht = 0.0f, hb0 = 0.0f, hb1 = 0.0f, hb2 = 0.0f;
    uint32_t hprim = PSY_NONE, hinst = PSY_NONE;
    bool done = false;

    uint64_t c_nodes = 0, c_mnodes = 0, c_tris = 0, c_mtris = 0, c_inst = 0;

    uint32_t cur = a.top_root;
    while (true) {
        if (cur == REF_EMPTY) {
            if (in_inst && sp == inst_base) {
                // Leave the instance: restore the top-level local ray
                // (tracer.rs:130-148).
                in_inst = false;
                lo = top_lo;
                di = top_di;
                oct = top_oct;
            }
            if (sp == 0) break;
            cur = STACK_POP();
            continue;
        }
        if (!(cur & REF_LEAF)) {
            // Internal node: test the four child boxes (bvh4.rs:120-169).
            const GpuNode* node = a.nodes + cur;
            const float4* nf = reinterpret_cast<const float4*>(node);
            const uint4 refs = *reinterpret_cast<const uint4*>(nf + 6);
            const uint4 tail = *reinterpret_cast<const uint4*>(nf + 7);
            const uint32_t meta = tail.z;
            const uint32_t cnt = meta & 7u;
            const uint32_t blen = (meta >> 8) & 0xffu;
            float4 lx, hx, ly, hy, lz, hz;
            if (blen <= 1) {
                lx = nf[0]; hx = nf[1]; ly = nf[2]; hy = nf[3]; lz = nf[4]; hz = nf[5];
                if (COUNT) c_nodes++;
            } else {
                uint32_t i1;
                float a2;
                const float4* m0;
                if (!lerp_slice_pos(blen, time, i1, a2)) {
                    m0 = reinterpret_cast<const float4*>(a.motion + tail.w + i1);
                    lx = m0[0]; hx = m0[1]; ly = m0[2]; hy = m0[3]; lz = m0[4]; hz = m0[5];
                } else {
                    m0 = reinterpret_cast<const float4*>(a.motion + tail.w + i1);
                    const float4* m1 = m0 + 6;
                    lx = lerp4(m0[0], m1[0], a2); hx = lerp4(m0[1], m1[1], a2);
                    ly = lerp4(m0[2], m1[2], a2); hy = lerp4(m0[3], m1[3], a2);
                    lz = lerp4(m0[4], m1[4], a2); hz = lerp4(m0[5], m1[5], a2);
                }
                if (COUNT) c_mnodes++;
            }
            uint32_t mask = 0;
            mask |= slab(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, lo, di, tmax) ? 1u : 0u;
            mask |= slab(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, lo, di, tmax) ? 2u : 0u;
            mask |= slab(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, lo, di, tmax) ? 4u : 0u;
            mask |= slab(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, lo, di, tmax) ? 8u : 0u;
            mask &= (1u << cnt) - 1u;

            // Visit order for this ray's octant; push all hit children but the
            // first in reverse order so the first is visited next.
            const uint32_t ord = ((oct < 4 ? tail.x : tail.y) >> (8u * (oct & 3u))) & 0xffu;
            uint32_t next = REF_EMPTY;
            for (int k = 3; k >= 0; --k) {
                if (k >= (int)cnt) continue;
                uint32_t c = (ord >> (2 * k)) & 3u;
                if (mask & (1u << c)) {
                    if (next != REF_EMPTY) STACK_PUSH(next);
                    next = c == 0 ? refs.x : (c == 1 ? refs.y : (c == 2 ? refs.z : refs.w));
                }
            }
            cur = next;
            continue;
        }

        const uint32_t first = (cur & ~REF_LEAF) >> 2;
        const uint32_t count = cur & 3u;
        if (!in_inst) {
            // Instance leaf (tracer.rs:95-128): move the ray into the
            // instance's space and descend into its mesh.
            const GpuInstance inst = a.instances[first];
            const GpuMesh mesh = a.meshes[inst.mesh];
            if (COUNT) c_inst++;
            if (inst.xf_count > 0) {
                uint32_t i1;
                float a2;
                Xf x;
                if (!lerp_slice_pos(inst.xf_count, time, i1, a2)) {
                    x = load_xf(a.xforms + inst.xf_start + i1);
                } else {
                    x = xf_lerp(load_xf(a.xforms + inst.xf_start + i1), load_xf(a.xforms + inst.xf_start + i1 + 1), a2);
                }
                lo = xf_point(x, o);
                V3 ld = xf_vector(x, d);
                di = v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z);
                oct = octant_of(di);
                // mat_space (triangle_mesh.rs:128-139, 184-189)
                space = inst.xf_count == 1 ? load_xf(a.inv_static + inst.inv_static) : xf_inverse(x);
                has_space = true;
            } else {
                has_space = false;
            }
            in_inst = true;
            inst_base = sp;
            inst_id = inst.id;
            tri_base = mesh.tri_offset;
            tri_stride = mesh.tri_stride;
            tri_samples = mesh.time_samples;
            cur = mesh.root;
            continue;
        }

        // Triangle leaf (triangle_mesh.rs:174-260).
        for (uint32_t j = 0; j < count; j++) {
            const uint8_t* rec = a.tris + tri_base + (uint64_t)(first + j) * tri_stride;
            V3 v0, v1, v2;
            uint32_t prim;
            if (tri_samples == 1) {
                const float4* f = reinterpret_cast<const float4*>(rec);
                float4 q0 = f[0], q1 = f[1], q2 = f[2];
                v0 = v3(q0.x, q0.y, q0.z);
                v1 = v3(q0.w, q1.x, q1.y);
                v2 = v3(q1.z, q1.w, q2.x);
                prim = __float_as_uint(q2.y);
                if (COUNT) c_tris++;
            } else {
                const float* f = reinterpret_cast<const float*>(rec);
                uint32_t i1;
                float a2;
                if (!lerp_slice_pos(tri_samples, time, i1, a2)) {
                    const float* s0 = f + i1 * 9;
                    v0 = v3(s0[0], s0[1], s0[2]);
                    v1 = v3(s0[3], s0[4], s0[5]);
                    v2 = v3(s0[6], s0[7], s0[8]);
                } else {
                    const float* s0 = f + i1 * 9;
                    const float* s1 = s0 + 9;
                    v0 = lerp3(v3(s0[0], s0[1], s0[2]), v3(s1[0], s1[1], s1[2]), a2);
                    v1 = lerp3(v3(s0[3], s0[4], s0[5]), v3(s1[3], s1[4], s1[5]), a2);
                    v2 = lerp3(v3(s0[6], s0[7], s0[8]), v3(s1[6], s1[7], s1[8]), a2);
                }
                prim = *reinterpret_cast<const uint32_t*>(f + 9 * tri_samples);
                if (COUNT) c_mtris++;
            }
            if (has_space) {
                v0 = xf_point(space, v0);
                v1 = xf_point(space, v1);
                v2 = xf_point(space, v2);
            }
            float t, b0, b1, b2;
            if (tri_test(o, pre, tmax, v0, v1, v2, t, b0, b1, b2)) {
                if (occl) {
                    kind = PSY_HIT_OCCLUDE;
                    done = true;
                    break;
                }
                tmax = t;
                kind = PSY_HIT_SURFACE;
                ht = t; hb0 = b0; hb1 = b1; hb2 = b2;
                hprim = prim;
                hinst = inst_id;
            }
        }
        if (done) break;
        cur = REF_EMPTY;
    }

    a.kind[ray] = kind;
    if (kind == PSY_HIT_SURFACE) {
        a.t[ray] = ht;
        a.b0[ray] = hb0;
        a.b1[ray] = hb1;
        a.b2[ray] = hb2;
        a.prim_id[ray] = hprim;
        a.inst_id[ray] = hinst;
    }
    if (COUNT) {
        atomicAdd(a.counters + 0, (unsigned long long)c_nodes);
        atomicAdd(a.counters + 1, (unsigned long long)c_mnodes);
        atomicAdd(a.counters + 2, (unsigned long long)c_tris);
        atomicAdd(a.counters + 3, (unsigned long long)c_mtris);
        atomicAdd(a.counters + 4, (unsigned long long)c_inst);
    }
}

}  // namespace

hipError_t launch_trace(const TraceArgs& args, bool counted, hipStream_t stream) {
    if (args.n == 0) return hipSuccess;
    uint64_t blocks = (args.n + BLOCK - 1) / BLOCK;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    if (counted)
        hipLaunchKernelGGL(trace_kernel<true>, dim3((uint32_t)blocks), dim3(BLOCK), 0, stream, args);
    else
        hipLaunchKernelGGL(trace_kernel<false>, dim3((uint32_t)blocks), dim3(BLOCK), 0, stream, args);
    return hipGetLastError();
}

}  // namespace psy
