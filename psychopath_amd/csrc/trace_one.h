// One ray through the instance and mesh BVH4s: the traversal of the
// one-lane-per-ray kernel (trace.hip) as a device function, also called by
// the single-launch renderer.  Each lane walks the instance BVH4 and, at an
// instance leaf, the mesh BVH4 in that instance's space: the same depth-first
// order the reference's ray-stream traversal gives every ray
// (src/accel/bvh4.rs:101-184, src/tracer.rs:87-150), the same box test
// (src/bbox4.rs:63-97), motion lerps (src/lerp.rs:20-35) and watertight
// triangle test in world space (src/surface/triangle.rs:58-149,
// src/surface/triangle_mesh.rs:174-260).  The pending-node stack keeps its
// first LDS_DEPTH entries in the lane's LDS column (stride `stride`) and the
// rest in a private array.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"
#include "psymath.h"
#include "trace.h"
#include "trace_device.h"

namespace psy {

struct TraceHit {
    uint8_t kind;
    float t, b0, b1, b2;
    uint32_t prim, inst;
};

struct TraceCount {
    uint64_t nodes = 0, mnodes = 0, tris = 0, mtris = 0, inst = 0;
};

// NESTED: false compiles the nested-assembly walk out (the caller
// guarantees a.groups is NULL); true follows a.groups at run time.
template <bool COUNT, int LDS_DEPTH, int SPILL_DEPTH, bool NESTED>
__device__ __forceinline__ TraceHit trace_one(const TraceArgs& a, V3 o, V3 d, float time, float tmax, bool occl,
                                               uint32_t* lds_col, int stride, TraceCount& ctr) {
#define T1_PUSH(v)                                                         \
    do {                                                                   \
        if (sp < LDS_DEPTH) lds_col[sp * stride] = (v);                    \
        else spill[sp - LDS_DEPTH] = (v);                                  \
        sp++;                                                              \
    } while (0)
#define T1_POP() (--sp, sp < LDS_DEPTH ? lds_col[sp * stride] : spill[sp - LDS_DEPTH])
    uint32_t spill[SPILL_DEPTH];
    int sp = 0;

    // Top-level local ray: update_local with the identity (tracer.rs:65-70).
    const Xf ident = xf_identity();
    const V3 top_lo = xf_point(ident, o);
    const V3 top_ld = xf_vector(ident, d);
    const V3 top_di = v3(1.0f / top_ld.x, 1.0f / top_ld.y, 1.0f / top_ld.z);
    const uint32_t top_oct = octant_of(top_di);
    const TriPre pre = tri_precompute(d);

    V3 lo = top_lo, di = top_di;
    uint32_t oct = top_oct;

    // Instance state.
    bool in_inst = false;
    int inst_base = 0;
    bool has_space = false;
    Xf space;
    uint32_t inst_id = PSY_NONE;
    uint64_t tri_base = 0;
    uint32_t tri_stride = 0, tri_samples = 1;
    // Nested scenes (a.groups set): the current assembly path and its first
    // instance record; level_ray sets the local ray of the path's level.
    uint32_t grp = 0, gbase = 0;
    auto level_ray = [&](const GpuGroup& g) {
        Xf x = xf_identity();
        if (g.xf_count > 0) {
            uint32_t i1;
            float a2;
            if (!lerp_slice_pos(g.xf_count, time, i1, a2)) x = load_xf(a.xforms + g.xf_start + i1);
            else x = xf_lerp(load_xf(a.xforms + g.xf_start + i1), load_xf(a.xforms + g.xf_start + i1 + 1), a2);
        }
        lo = xf_point(x, o);
        const V3 ld = xf_vector(x, d);
        di = v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z);
        oct = octant_of(di);
    };

    TraceHit h{PSY_HIT_MISS, 0.0f, 0.0f, 0.0f, 0.0f, PSY_NONE, PSY_NONE};
    bool done = false;

    uint32_t cur = a.top_root;
    while (true) {
        if (cur == REF_EMPTY) {
            if (in_inst && sp == inst_base) {
                // Leave the instance: restore the local ray of its assembly
                // level (tracer.rs:130-148).
                in_inst = false;
                if (NESTED && a.groups) {
                    level_ray(a.groups[grp]);
                } else {
                    lo = top_lo;
                    di = top_di;
                    oct = top_oct;
                }
            }
            if (sp == 0) break;
            cur = T1_POP();
            continue;
        }
        if (!(cur & REF_LEAF)) {
            // Internal node: test the four child boxes (bvh4.rs:120-169).
            const GpuNode* node = a.nodes + cur;
            const float4* nf = reinterpret_cast<const float4*>(node);
            uint4 refs = *reinterpret_cast<const uint4*>(nf + 6);
            const uint4 tail = *reinterpret_cast<const uint4*>(nf + 7);
            refs.x &= ~REF_CODE_MASK;  // the traversal code's bits (layout.h)
            refs.y &= ~REF_CODE_MASK;
            const uint32_t meta = tail.z;
            const uint32_t cnt = meta & 7u;
            const uint32_t blen = (meta >> 8) & 0xffu;
            float4 lx, hx, ly, hy, lz, hz;
            if (blen <= 1) {
                lx = nf[0]; hx = nf[1]; ly = nf[2]; hy = nf[3]; lz = nf[4]; hz = nf[5];
                if (COUNT) ctr.nodes++;
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
                if (COUNT) ctr.mnodes++;
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
                    if (next != REF_EMPTY) T1_PUSH(next);
                    next = c == 0 ? refs.x : (c == 1 ? refs.y : (c == 2 ? refs.z : refs.w));
                }
            }
            cur = next;
            continue;
        }

        const uint32_t first = (cur & ~REF_LEAF) >> 2;
        const uint32_t count = cur & 3u;
        const uint32_t ii = gbase + first;
        if (NESTED && a.groups && !in_inst && count == 0u) {
            // Level marker: the sub-assembly is done; back to the enclosing
            // assembly's path and local ray (tracer.rs:131-148).
            grp = first;
            const GpuGroup g = a.groups[grp];
            gbase = g.base;
            level_ray(g);
            cur = REF_EMPTY;
            continue;
        }
        if (NESTED && a.groups && !in_inst && a.instances[ii].kind == PSY_OBJECT_ASSEMBLY) {
            // Sub-assembly instance (tracer.rs:95-128): remember the
            // enclosing path on the stack and walk the sub-assembly's
            // instance BVH in its space.
            if (COUNT) ctr.inst++;
            T1_PUSH(make_level_marker(grp));
            grp = a.instances[ii].mesh;
            const GpuGroup g = a.groups[grp];
            gbase = g.base;
            level_ray(g);
            cur = g.root;
            continue;
        }
        if (!in_inst && a.instances[ii].kind != PSY_OBJECT_MESH) {
            // Light instance: test the light's surface in place.
            const GpuInstance inst = a.instances[ii];
            if (COUNT) ctr.inst++;
            const Xf x = instance_xf(a.xforms, inst, time);
            float lt, l0, l1, l2;
            uint32_t lp;
            if (light_leaf_test(a.rect_lights, a.sphere_lights, a.light_samples, time, inst, x, o, pre,
                                xf_point(x, o), xf_vector(x, d), tmax, lt, l0, l1, l2, lp)) {
                if (occl) {
                    h.kind = PSY_HIT_OCCLUDE;
                    break;
                }
                tmax = lt;
                h = TraceHit{PSY_HIT_SURFACE, lt, l0, l1, l2, lp, inst.id};
            }
            cur = REF_EMPTY;
            continue;
        }
        if (!in_inst) {
            // Instance leaf (tracer.rs:95-128): move the ray into the
            // instance's space and descend into its mesh.
            const GpuInstance inst = a.instances[ii];
            const GpuMesh mesh = a.meshes[inst.mesh];
            if (COUNT) ctr.inst++;
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
                if (COUNT) ctr.tris++;
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
                if (COUNT) ctr.mtris++;
            }
            if (has_space) {
                v0 = xf_point(space, v0);
                v1 = xf_point(space, v1);
                v2 = xf_point(space, v2);
            }
            float t, b0, b1, b2;
            if (tri_test(o, pre, tmax, v0, v1, v2, t, b0, b1, b2)) {
                if (occl) {
                    h.kind = PSY_HIT_OCCLUDE;
                    done = true;
                    break;
                }
                tmax = t;
                h = TraceHit{PSY_HIT_SURFACE, t, b0, b1, b2, prim, inst_id};
            }
        }
        if (done) break;
        cur = REF_EMPTY;
    }

    return h;
#undef T1_PUSH
#undef T1_POP
}

}  // namespace psy
