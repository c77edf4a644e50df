// Persistent BVH4 traversal for gfx950 with per-lane ray refill.
//
// Same traversal as trace.hip (each ray walks the reference's depth-first
// order, src/accel/bvh4.rs:101-184, with the same box and triangle tests),
// reorganised for throughput on incoherent rays:
//   * one wave64 per workgroup, launched only as many waves as the chip holds
//     at once; each wave keeps taking chunks of rays from a global counter;
//   * lanes whose ray finished are refilled in bulk once at least
//     `refill_min` lanes of the wave are idle, from the wave's chunk (an
//     atomic only when the chunk runs short), the new rays spread over the
//     idle lanes by a ballot prefix count;
//   * the traversal is "while-while": lanes descend internal nodes (popping
//     past dead ends) until the whole wave stands on leaves, then run the
//     leaf tests together;
//   * kernels are specialised on whether the scene has instance transforms
//     (XF), time-sampled bounds or vertices (MOTION) and nested assemblies
//     (NESTED: each assembly's own instance BVH walked in its space, as
//     trace_assembly recurses, tracer.rs:87-150), so the common static case
//     keeps no transform, lerp or nesting state in registers.
#include <hip/hip_runtime.h>

#include "layout.h"
#include "psymath.h"
#include "trace.h"
#include "trace_device.h"

namespace psy {

namespace {

constexpr int WAVE = 64;
// Most rays a wave takes from the global counter at once.  Larger chunks
// spread the rays in flight over a wider stretch of the queue: on queues in
// pixel order (a renderer's camera and bounce rays) that costs coherence, 4%
// on the large instanced scene at 512, while 64 leaves the counter contended
// on small scenes.
constexpr uint32_t CHUNK_MAX = 256;

// Short stack: the newest LDS_DEPTH entries of each lane live in an LDS ring
// (one column per lane, stride WAVE); when the ring is full the oldest entry
// moves to a private overflow array and comes back only when the stack
// unwinds to it, so the pushes and pops of the hot path touch LDS only.
// `sp` counts entries, entries [0, bot) are in the overflow array.  Macros
// keep sp and bot in registers (a struct holding the array would live in
// private memory as a whole).
#define SSTACK_PUSH(v)                                                    \
    do {                                                                  \
        if (sp - bot == LDS_DEPTH) {                                      \
            spill[bot] = lds_col[(bot % LDS_DEPTH) * WAVE];               \
            bot++;                                                        \
        }                                                                 \
        lds_col[(sp % LDS_DEPTH) * WAVE] = (v);                           \
        sp++;                                                             \
    } while (0)
#define SSTACK_POP(out)                                                   \
    do {                                                                  \
        sp--;                                                             \
        if (sp < bot) {                                                   \
            bot = sp;                                                     \
            (out) = spill[sp];                                            \
        } else {                                                          \
            (out) = lds_col[(sp % LDS_DEPTH) * WAVE];                     \
        }                                                                 \
    } while (0)

// Node parts a static descent step reads: the six box rows and the child
// references (the last 16 bytes are not needed, layout.h).
constexpr uint32_t NODE_PARTS = 7;

// The local ray of an assembly level: the lerped composed transforms into
// the assembly applied to the world ray, or the identity at the root
// (update_local, tracer.rs:65-70, 101-109, 131-148).
__device__ __forceinline__ void level_local_ray(const TraceArgs& a, const GpuGroup& g, float time, V3 o, V3 d, V3& lo,
                                                V3& di, uint32_t& oct) {
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
}

template <bool COUNT, bool XF, bool MOTION, int LDS_DEPTH, int MIN_WAVES, bool COOP, bool NESTED>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(MIN_WAVES)))
void trace_persistent_kernel(TraceArgs a, unsigned int* work, uint32_t refill_min, uint32_t leaf_min) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * WAVE];
    // COOP: the nodes the wave's lanes stand on are fetched by the wave
    // together, consecutive lanes loading consecutive 16-byte parts of one
    // node (one cache line per node instead of one tag lookup per lane and
    // part), staged here and read back by their lanes.
    __shared__ uint4 lds_nodes[COOP ? WAVE * NODE_PARTS : 1];
    __shared__ uint32_t lds_slot[COOP ? WAVE : 1];
    // TRAVERSAL_TABLE (8 octants x 48 codes): static nodes carry only their
    // code, in spare bits of their child references (layout.h).
    __shared__ uint32_t lds_order[8 * 48 / 4];
    const int lane = threadIdx.x;
    uint32_t* const lds_col = lds_stack + lane;
    if (!MOTION) {
        const uint32_t* tab = reinterpret_cast<const uint32_t*>(a.order_table);
        for (int i = lane; i < 8 * 48 / 4; i += WAVE) lds_order[i] = tab[i];
        __syncthreads();
    }
    const uint8_t* const order_lds = reinterpret_cast<const uint8_t*>(lds_order);
    uint32_t spill[TRACE_STACK_CAPACITY];
    int sp = 0, bot = 0;

    const uint32_t n = a.n_dev ? *a.n_dev : (uint32_t)a.n;
    const Xf ident = xf_identity();

    // Per-lane ray state.
    bool has = false;
    uint32_t ray = 0;
    V3 o{}, d{};
    float time = 0.0f, tmax = 0.0f;
    bool occl = false;
    TriPre pre{};
    V3 lo{}, di{};
    uint32_t oct = 0;
    uint32_t cur = REF_EMPTY;
    bool in_inst = false;
    int inst_base = 0;
    bool has_space = false;
    Xf space;
    uint32_t inst_id = PSY_NONE;
    uint32_t grp = 0, gbase = 0;  // NESTED: the current assembly path and its first instance record
    uint64_t tri_base = 0;
    uint32_t tri_stride = sizeof(GpuTriStatic), tri_samples = 1;
    uint8_t kind = PSY_HIT_MISS;
    float ht = 0.0f, hb0 = 0.0f, hb1 = 0.0f, hb2 = 0.0f;
    uint32_t hprim = PSY_NONE, hinst = PSY_NONE;
    uint64_t c_nodes = 0, c_mnodes = 0, c_tris = 0, c_mtris = 0, c_inst = 0;
    bool exhausted = false;
    // The wave's chunk: rays [next, end) taken from the counter and not yet
    // handed to a lane; `seen` is where the counter stood after the wave's
    // last fetch, and `drained` that it had passed n then.
    uint32_t next = 0, end = 0, seen = 0;
    bool drained = false;

    for (;;) {
        // ---- refill idle lanes ------------------------------------------
        // Idle lanes take the rest of the wave's chunk first; one atomic
        // fetches a new chunk only when the rest is short, sized by guided
        // self-scheduling (the remaining rays over 2x the waves, at most
        // CHUNK_MAX), so the counter sees few atomics while the last rays
        // still spread over all waves.
        const uint64_t idle = __ballot(!has);
        if (!exhausted && (uint32_t)__popcll(idle) >= refill_min) {
            const uint32_t want = (uint32_t)__popcll(idle);
            const uint32_t avail = end - next;
            uint32_t fresh = 0, fresh_end = 0;
            if (avail < want && !drained) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                const uint32_t left = n > seen ? n - seen : 0u;
                const uint32_t take =
                    max(want - avail, min(left / (2u * (uint32_t)gridDim.x), (uint32_t)CHUNK_MAX));
                if (lane == leader) fresh = atomicAdd(work, take);
                fresh = __shfl(fresh, leader);
                seen = fresh + take;
                drained = seen >= n;
                fresh_end = fresh < n ? min(seen, n) : fresh;
            }
            if (!has) {
                const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                const uint32_t mine = rank < avail ? next + rank : fresh + (rank - avail);
                if (rank < avail || mine < fresh_end) {
                    has = true;
                    ray = mine;
                    o = v3(a.ox[ray], a.oy[ray], a.oz[ray]);
                    d = v3(a.dx[ray], a.dy[ray], a.dz[ray]);
                    time = MOTION || XF ? a.time[ray] : 0.0f;
                    tmax = a.max_t[ray];
                    occl = a.all_occlusion != 0 || (a.flags != nullptr && (a.flags[ray] & 1u) != 0);
                    // Top-level local ray: update_local with the identity
                    // (tracer.rs:65-70).
                    lo = xf_point(ident, o);
                    V3 ld = xf_vector(ident, d);
                    di = v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z);
                    oct = octant_of(di);
                    pre = tri_precompute(d);
                    cur = a.top_root;
                    sp = 0;
                    bot = 0;
                    in_inst = false;
                    grp = 0;
                    gbase = 0;
                    kind = PSY_HIT_MISS;
                }
            }
            if (avail >= want) {
                next += want;
            } else {
                // The lanes took the whole rest and the head of the new chunk
                // (none when the counter had already passed n).
                next = min(fresh + (want - avail), fresh_end);
                end = fresh_end;
            }
            exhausted = drained && next == end;
        }
        if (__ballot(has) == 0) break;
        if (!has) continue;

        // ---- descend internal nodes ---------------------------------------
        // leaf_min 0: each lane descends until it stands on a leaf or on a
        // dead end (a node whose children all miss), which pops after the
        // leaf round.  leaf_min 1..64: dead ends pop inside the loop, so the
        // lane keeps working while others descend, and the loop runs while
        // some lane stands on an internal node and fewer than `leaf_min`
        // lanes wait (at a leaf, or at the end of their level: the
        // instance's subtree or the whole stack); lanes still on internal
        // nodes then carry on after the leaf round.  Every lane takes its own
        // depth-first steps in the same order either way; only their timing
        // changes, so results do not depend on leaf_min.
        //
        // One descent step of a lane standing on an internal node: its four
        // child boxes, then the hit children pushed in visit order.
        // The wave's lanes standing on internal nodes (`inner`, ballot `m`)
        // have their nodes' parts loaded by all active lanes together:
        // parts item = slot * 7 + part, item k of a batch by the k-th active
        // lane, four loads in flight per lane; returns this lane's slot.
        auto coop_fetch = [&](bool inner, uint64_t m) -> uint32_t {
            const uint64_t below = (1ull << lane) - 1ull;
            const uint32_t slot = (uint32_t)__popcll(m & below);
            const uint32_t items = NODE_PARTS * (uint32_t)__popcll(m);
            const uint64_t act = __ballot(1);
            const uint32_t na = (uint32_t)__popcll(act), ar = (uint32_t)__popcll(act & below);
            if (inner) lds_slot[slot] = cur;
            __syncthreads();
            for (uint32_t base = 0; base < items; base += 4 * na) {
                uint4 v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t item = base + (uint32_t)k * na + ar;
                    if (item < items)
                        v[k] = reinterpret_cast<const uint4*>(a.nodes + lds_slot[item / NODE_PARTS])[item % NODE_PARTS];
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t item = base + (uint32_t)k * na + ar;
                    if (item < items) lds_nodes[item] = v[k];
                }
            }
            __syncthreads();
            return slot;
        };
        auto node_step = [&](uint32_t slot) {
                const float4* nf = reinterpret_cast<const float4*>(a.nodes + cur);
                const uint4* staged = lds_nodes + (COOP ? slot * NODE_PARTS : 0u);
                uint4 refs = COOP ? staged[6] : *reinterpret_cast<const uint4*>(nf + 6);
                // Static scenes read seven of the node's eight 16-byte parts:
                // the child count follows from the absent references and the
                // visit order from the traversal code in the references' spare
                // bits; time-sampled scenes read the last part for the bounds'
                // sample count and location.
                uint4 tail = make_uint4(0u, 0u, 0u, 0u);
                if (MOTION) tail = *reinterpret_cast<const uint4*>(nf + 7);
                const uint32_t code = ((refs.x >> REF_CODE_SHIFT) & 7u) | (((refs.y >> REF_CODE_SHIFT) & 7u) << 3);
                refs.x &= ~REF_CODE_MASK;
                refs.y &= ~REF_CODE_MASK;
                const uint32_t cnt =
                    MOTION ? (tail.z & 7u) : 2u + (refs.z != REF_EMPTY ? 1u : 0u) + (refs.w != REF_EMPTY ? 1u : 0u);
                float4 lx, hx, ly, hy, lz, hz;
                const uint32_t blen = MOTION ? ((tail.z >> 8) & 0xffu) : 1u;
                if (COOP) {
                    const float4* sf = reinterpret_cast<const float4*>(staged);
                    lx = sf[0]; hx = sf[1]; ly = sf[2]; hy = sf[3]; lz = sf[4]; hz = sf[5];
                } else if (!MOTION || blen <= 1) {
                    lx = nf[0]; hx = nf[1]; ly = nf[2]; hy = nf[3]; lz = nf[4]; hz = nf[5];
                    if (COUNT) c_nodes++;
                } else {
                    uint32_t i1;
                    float a2;
                    const float4* m0 = reinterpret_cast<const float4*>(a.motion + tail.w);
                    if (!lerp_slice_pos(blen, time, i1, a2)) {
                        m0 += 6 * i1;
                        lx = m0[0]; hx = m0[1]; ly = m0[2]; hy = m0[3]; lz = m0[4]; hz = m0[5];
                    } else {
                        m0 += 6 * i1;
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
                const uint32_t ord = MOTION ? ((oct < 4 ? tail.x : tail.y) >> (8u * (oct & 3u))) & 0xffu
                                            : (uint32_t)order_lds[oct * 48u + code];
                uint32_t next = REF_EMPTY;
                for (int k = 3; k >= 0; --k) {
                    if (k >= (int)cnt) continue;
                    const uint32_t c = (ord >> (2 * k)) & 3u;
                    if (mask & (1u << c)) {
                        if (next != REF_EMPTY) SSTACK_PUSH(next);
                        next = c == 0 ? refs.x : (c == 1 ? refs.y : (c == 2 ? refs.z : refs.w));
                    }
                }
                cur = next;
        };
        if (leaf_min == 0) {
            // Each lane descends until it stands on a leaf or a dead end.
            if (COOP) {
                for (;;) {
                    const bool inner = cur != REF_EMPTY && !(cur & REF_LEAF);
                    const uint64_t m = __ballot(inner);
                    if (m == 0) break;
                    const uint32_t slot = coop_fetch(inner, m);
                    if (inner) node_step(slot);
                }
            } else {
                while (cur != REF_EMPTY && !(cur & REF_LEAF)) node_step(0u);
            }
        } else {
            const uint32_t n_active = (uint32_t)__popcll(__ballot(1));
            for (;;) {
                if (cur == REF_EMPTY && sp > (in_inst ? inst_base : 0)) SSTACK_POP(cur);
                const bool inner = cur != REF_EMPTY && !(cur & REF_LEAF);
                const uint64_t m = __ballot(inner);
                const uint32_t n_inner = (uint32_t)__popcll(m);
                if (n_inner == 0 || n_active - n_inner >= leaf_min) break;
                const uint32_t slot = COOP ? coop_fetch(inner, m) : 0u;
                if (inner) node_step(slot);
            }
        }

        // ---- leaf --------------------------------------------------------
        bool finished = false;
        if (cur != REF_EMPTY && (cur & REF_LEAF)) {
            const uint32_t first = (cur & ~REF_LEAF) >> 2;
            const uint32_t count = cur & 3u;
            const uint32_t ii = NESTED ? gbase + first : first;
            if (NESTED && !in_inst && count == 0u) {
                // Level marker: the sub-assembly is done; back to the
                // enclosing assembly's path and local ray (tracer.rs:131-148).
                grp = first;
                const GpuGroup g = a.groups[grp];
                gbase = g.base;
                level_local_ray(a, g, time, o, d, lo, di, oct);
                cur = REF_EMPTY;
            } else if (NESTED && !in_inst && a.instances[ii].kind == PSY_OBJECT_ASSEMBLY) {
                // Sub-assembly instance (tracer.rs:95-128): remember the
                // enclosing path on the stack, move the ray into the
                // sub-assembly's space and walk its instance BVH.
                if (COUNT) c_inst++;
                SSTACK_PUSH(make_level_marker(grp));
                grp = a.instances[ii].mesh;
                const GpuGroup g = a.groups[grp];
                gbase = g.base;
                level_local_ray(a, g, time, o, d, lo, di, oct);
                cur = g.root;
            } else if (!in_inst && XF && a.instances[ii].kind != PSY_OBJECT_MESH) {
                // Light instance (tracer.rs:95-128 with a light object; light
                // instances always carry transforms).
                const GpuInstance inst = a.instances[ii];
                if (COUNT) c_inst++;
                const Xf x = instance_xf(a.xforms, inst, time);
                const V3 llo = xf_point(x, o);
                const V3 lld = xf_vector(x, d);
                float lt, l0, l1, l2;
                uint32_t lp;
                if (light_leaf_test(a.rect_lights, a.sphere_lights, a.light_samples, time, inst, x, o, pre, llo, lld,
                                    tmax, lt, l0, l1, l2, lp)) {
                    if (occl) {
                        kind = PSY_HIT_OCCLUDE;
                        finished = true;
                    } else {
                        tmax = lt;
                        kind = PSY_HIT_SURFACE;
                        ht = lt; hb0 = l0; hb1 = l1; hb2 = l2;
                        hprim = lp;
                        hinst = inst.id;
                    }
                }
                cur = REF_EMPTY;
            } else if (!in_inst) {
                // Instance leaf (tracer.rs:95-128).
                const GpuInstance inst = a.instances[ii];
                const GpuMesh mesh = a.meshes[inst.mesh];
                if (COUNT) c_inst++;
                if (XF) {
                    if (inst.xf_count > 0) {
                        uint32_t i1;
                        float a2;
                        Xf x;
                        if (!lerp_slice_pos(inst.xf_count, time, i1, a2))
                            x = load_xf(a.xforms + inst.xf_start + i1);
                        else
                            x = xf_lerp(load_xf(a.xforms + inst.xf_start + i1),
                                        load_xf(a.xforms + inst.xf_start + i1 + 1), a2);
                        lo = xf_point(x, o);
                        V3 ld = xf_vector(x, d);
                        di = v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z);
                        oct = octant_of(di);
                        space = inst.xf_count == 1 ? load_xf(a.inv_static + inst.inv_static) : xf_inverse(x);
                        has_space = true;
                    } else {
                        has_space = false;
                    }
                }
                in_inst = true;
                inst_base = sp;
                inst_id = inst.id;
                tri_base = mesh.tri_offset;
                if (MOTION) {
                    tri_stride = mesh.tri_stride;
                    tri_samples = mesh.time_samples;
                }
                cur = mesh.root;
            } else {
                // Triangle leaf (triangle_mesh.rs:174-260).
                for (uint32_t j = 0; j < count; j++) {
                    const uint8_t* rec = a.tris + tri_base + (uint64_t)(first + j) * tri_stride;
                    V3 v0, v1, v2;
                    uint32_t prim;
                    if (!MOTION || tri_samples == 1) {
                        const float4* f = reinterpret_cast<const float4*>(rec);
                        const float4 q0 = f[0], q1 = f[1], q2 = f[2];
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
                    if (XF && has_space) {
                        v0 = xf_point(space, v0);
                        v1 = xf_point(space, v1);
                        v2 = xf_point(space, v2);
                    }
                    float t, b0, b1, b2;
                    if (tri_test(o, pre, tmax, v0, v1, v2, t, b0, b1, b2)) {
                        if (occl) {
                            kind = PSY_HIT_OCCLUDE;
                            finished = true;
                            break;
                        }
                        tmax = t;
                        kind = PSY_HIT_SURFACE;
                        ht = t; hb0 = b0; hb1 = b1; hb2 = b2;
                        hprim = prim;
                        hinst = inst_id;
                    }
                }
                cur = REF_EMPTY;
            }
        }

        // ---- pop, leaving the instance when its subtree is done -----------
        if (!finished && cur == REF_EMPTY) {
            if (in_inst && sp == inst_base) {
                in_inst = false;
                if (NESTED) {
                    level_local_ray(a, a.groups[grp], time, o, d, lo, di, oct);
                } else if (XF) {
                    lo = xf_point(ident, o);
                    V3 ld = xf_vector(ident, d);
                    di = v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z);
                    oct = octant_of(di);
                }
            }
            if (sp == 0) finished = true;
            else SSTACK_POP(cur);
        }

        if (finished) {
            if (a.rec != nullptr) {
                // One 32-byte record per ray, two 16-byte stores: rays finish
                // out of order, and a whole record fills a memory sector
                // where the seven arrays of the SoA form leave seven partial
                // lines.
                const bool s = kind == PSY_HIT_SURFACE;
                uint4* r = a.rec + 2ull * ray;
                r[0] = make_uint4(s ? __float_as_uint(ht) : 0u, s ? __float_as_uint(hb0) : 0u,
                                  s ? __float_as_uint(hb1) : 0u, s ? __float_as_uint(hb2) : 0u);
                r[1] = make_uint4(s ? hprim : PSY_NONE, s ? hinst : PSY_NONE, (uint32_t)kind, 0u);
            } else {
                a.kind[ray] = kind;
                if (kind == PSY_HIT_SURFACE) {
                    a.t[ray] = ht;
                    a.b0[ray] = hb0;
                    a.b1[ray] = hb1;
                    a.b2[ray] = hb2;
                    a.prim_id[ray] = hprim;
                    a.inst_id[ray] = hinst;
                }
            }
            has = false;
        }
    }
    if (COUNT) {
        atomicAdd(a.counters + 0, (unsigned long long)c_nodes);
        atomicAdd(a.counters + 1, (unsigned long long)c_mnodes);
        atomicAdd(a.counters + 2, (unsigned long long)c_tris);
        atomicAdd(a.counters + 3, (unsigned long long)c_mtris);
        atomicAdd(a.counters + 4, (unsigned long long)c_inst);
    }
}

template <bool COUNT, bool XF, bool MOTION, int D, bool NESTED>
hipError_t launch_one(const TraceArgs& a, unsigned int* work, const TraceTuning& tune, hipStream_t s) {
    // Register budget: 8 waves per SIMD caps the kernel at 64 VGPRs (the
    // compiler keeps the rest in private memory), 4 lets it use up to 128.
    // The cooperative node fetch serves static single-level scenes.
    if constexpr (!COUNT && !MOTION && !NESTED) {
        if (tune.coop) {
            if (tune.occupancy >= 8)
                hipLaunchKernelGGL((trace_persistent_kernel<false, XF, false, D, 8, true, false>), dim3(tune.waves),
                                   dim3(WAVE), 0, s, a, work, tune.refill_min, tune.leaf_min);
            else
                hipLaunchKernelGGL((trace_persistent_kernel<false, XF, false, D, 4, true, false>), dim3(tune.waves),
                                   dim3(WAVE), 0, s, a, work, tune.refill_min, tune.leaf_min);
            return hipGetLastError();
        }
    }
    if (!COUNT && tune.occupancy >= 8)
        hipLaunchKernelGGL((trace_persistent_kernel<COUNT, XF, MOTION, D, 8, false, NESTED>), dim3(tune.waves), dim3(WAVE), 0,
                           s, a, work, tune.refill_min, tune.leaf_min);
    else
        hipLaunchKernelGGL((trace_persistent_kernel<COUNT, XF, MOTION, D, 4, false, NESTED>), dim3(tune.waves), dim3(WAVE), 0,
                           s, a, work, tune.refill_min, tune.leaf_min);
    return hipGetLastError();
}

template <bool COUNT, bool XF, bool MOTION, bool NESTED>
hipError_t launch_depth(const TraceArgs& a, unsigned int* work, const TraceTuning& tune, hipStream_t s) {
    if (COUNT || NESTED) return launch_one<COUNT, XF, MOTION, 16, NESTED>(a, work, tune, s);
    switch (tune.lds_depth) {
        case 8: return launch_one<COUNT, XF, MOTION, 8, false>(a, work, tune, s);
        case 24: return launch_one<COUNT, XF, MOTION, 24, false>(a, work, tune, s);
        default: return launch_one<COUNT, XF, MOTION, 16, false>(a, work, tune, s);
    }
}

template <bool COUNT>
hipError_t launch_feat(const TraceArgs& a, unsigned int* work, const TraceTuning& tune, bool xf, bool motion,
                       hipStream_t s) {
    // Nested scenes take the transform path, with or without time samples.
    if (a.groups) {
        if (motion) return launch_depth<COUNT, true, true, true>(a, work, tune, s);
        return launch_depth<COUNT, true, false, true>(a, work, tune, s);
    }
    if (xf && motion) return launch_depth<COUNT, true, true, false>(a, work, tune, s);
    if (xf) return launch_depth<COUNT, true, false, false>(a, work, tune, s);
    if (motion) return launch_depth<COUNT, false, true, false>(a, work, tune, s);
    return launch_depth<COUNT, false, false, false>(a, work, tune, s);
}

}  // namespace

// Sets *work = 0 with a kernel, so the launch sequence can be captured in a
// graph (no memset node).
__global__ void reset_work_kernel(unsigned int* work) { *work = 0u; }

hipError_t launch_trace_persistent(const TraceArgs& args, bool counted, bool has_xf, bool has_motion,
                                   unsigned int* work, const TraceTuning& tune, hipStream_t stream) {
    if (args.n == 0) return hipSuccess;
    if (args.n >= 0x80000000ull || tune.waves == 0 || tune.refill_min == 0 || tune.refill_min > 64 ||
        tune.leaf_min > 64)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_work_kernel, dim3(1), dim3(1), 0, stream, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return counted ? launch_feat<true>(args, work, tune, has_xf, has_motion, stream)
                   : launch_feat<false>(args, work, tune, has_xf, has_motion, stream);
}

}  // namespace psy
