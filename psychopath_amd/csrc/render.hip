// Wavefront path tracer for gfx950 (include/psy_render.h).
//
// The reference renders buckets of pixels per CPU thread, advancing every
// light path of a bucket one event at a time (LightPath::next,
// src/renderer.rs:449-692) between calls to Tracer::trace.  Here the paths of
// a whole wave (up to millions) live in HBM as structure-of-arrays and each
// event is one kernel over the wave:
//
//   camera    Camera::generate_ray + LightPath::new for every path slot
//   trace     closest-hit queue and shadow queue through the persistent
//             traversal kernel (counts read on the device)
//   shadow    unoccluded shadow rays add their pending light
//   sort      closest hits binned by closure kind (miss, light, Lambert,
//             GGX, emit) so the shading waves run one closure's code
//   shade     emission / background, next-event estimation through the light
//             tree, BSDF sampling; appends the next shadow and bounce rays
//   film      path colour -> XYZ, added per pixel in sample order
//
// A path's shadow ray and its next bounce ray are traced in the same step
// (the reference traces them one after the other); the colour additions keep
// the reference's order per path, so every path computes what the
// reference's does.  The steps of one wave are captured once into a hipGraph
// and replayed per wave; a device-side wave counter selects each replay's
// plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "color_host.h"
#include "render_internal.h"
#include "trace_one.h"

using namespace psy;
using namespace psy::api;
using namespace psy::rd;
using namespace psy::render;

namespace {

__device__ __forceinline__ const WaveSeg& find_seg(const WavePlan& p, uint32_t i) {
    uint32_t s = 0;
    while (s + 1 < p.nseg && i >= p.seg[s + 1].slot0) s++;
    return p.seg[s];
}

// Block-aggregated appends to two queues: one atomic per block and queue,
// shared by the block's KBLOCK / 64 waves (`cnt` and `base` are the block's
// LDS, 2 x KBLOCK / 64 words each).  Every thread of the block calls it.
__device__ __forceinline__ void block_append2(uint32_t* cs, bool ws, uint32_t* cb, bool wb, uint32_t* cnt,
                                              uint32_t* base, uint32_t& at_s, uint32_t& at_b) {
    constexpr int NW = KBLOCK / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t ms = __ballot(ws), mb = __ballot(wb);
    if (lane == 0) {
        cnt[w] = (uint32_t)__popcll(ms);
        cnt[NW + w] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        uint32_t tot = 0;
        for (int j = 0; j < NW; j++) tot += cnt[k * NW + j];
        uint32_t b0 = tot ? atomicAdd(k == 0 ? cs : cb, tot) : 0u;
        for (int j = 0; j < NW; j++) {
            base[k * NW + j] = b0;
            b0 += cnt[k * NW + j];
        }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    at_s = base[w] + (uint32_t)__popcll(ms & below);
    at_b = base[NW + w] + (uint32_t)__popcll(mb & below);
}

__device__ __forceinline__ void queue_put(const Queue& q, uint32_t at, uint32_t path, V3 o, V3 d, float time,
                                          float max_t) {
    q.path[at] = path;
    q.ox[at] = o.x; q.oy[at] = o.y; q.oz[at] = o.z;
    q.dx[at] = d.x; q.dy[at] = d.y; q.dz[at] = d.z;
    q.time[at] = time;
    q.max_t[at] = max_t;
}

// ---- kernels -------------------------------------------------------------------

__global__ void wave_begin_kernel(uint32_t* wave_idx, uint32_t value) { *wave_idx = value; }
__global__ void wave_advance_kernel(uint32_t* wave_idx) { *wave_idx += 1u; }

// Zeroes the closest and shadow output counters and the sort bins.
__global__ void reset_counts_kernel(uint32_t* a, uint32_t* b, uint32_t* bins) {
    const int i = threadIdx.x;
    if (i == 0) *a = 0u;
    if (i == 1) *b = 0u;
    if (i < 2 * (int)NUM_BINS) bins[i] = 0u;
}

__global__ void zero_film_kernel(float* film, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        film[i] = 0.0f;
}

// The state of one light path (LightPath, renderer.rs:362-447) while it is
// being advanced: the wavefront kernels keep it in HBM between events, the
// single-launch renderer in registers.
struct PathRegs {
    S4 color, atten, pending;
    float csp;  // closure_sample_pdf
    float wl, time;
    uint32_t seed, bounces, x, y, sample;
    // The last bounce ray sampled and its filter (next_bounce_ray,
    // next_attenuation_fac).  The reference keeps them when a later BSDF
    // sample fails, and a path that also found a light then continues along
    // this earlier ray after its shadow ray (renderer.rs:599-646, 670-690).
    bool has_next;
    V3 next_o, next_d;
    S4 next_fac;
};

// lerp_slice of a lens list (camera.rs:76-79).
__device__ __forceinline__ float lens_lerp(const float* s, uint32_t n, float time) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(n, time, i1, a2)) return s[i1];
    return lerpf(s[i1], s[i1 + 1], a2);
}

// LightPath::new and Camera::generate_ray for path g of a request
// (renderer.rs:233-289, camera.rs:74-102).
__device__ __forceinline__ void camera_path(const ShadeScene& sc, const ReqParams& rp, uint64_t g, PathRegs& ps,
                                            V3& o, V3& d) {
    const uint32_t pixel = (uint32_t)(g / rp.spp);
    const uint32_t si = (uint32_t)(g % rp.spp);
    const uint32_t x = rp.rx + pixel % rp.rw, y = rp.ry + pixel / rp.rw;
    const S4 d0 = sample_4d(sc.hdims, sc.hperms, si, 0, x, y, rp.seed);
    const S4 d1 = sample_4d(sc.hdims, sc.hperms, si, 1, x, y, rp.seed);
    const float filter_x = probit(d0.v[3], 2.0f / 6.0f) + 0.5f;
    const float filter_y = probit(d1.v[0], 2.0f / 6.0f) + 0.5f;
    const float samp_x = (filter_x + (float)x) * rp.cmpx;
    const float samp_y = (filter_y + (float)y) * rp.cmpy;
    const float img_x = (samp_x - 0.5f) * rp.x_extent;
    const float img_y = (0.5f - samp_y) * rp.y_extent;
    const float time = d0.v[0];
    const float wl = map_0_1_to_wavelength(golden_ratio_sample(si, hash_u32((x << 16) ^ y, rp.seed)));
    Xf xf;
    {
        uint32_t i1;
        float a2;
        if (!lerp_slice_pos(rp.cam_xf_count, time, i1, a2)) xf = load_xf(rp.cam_xf + i1);
        else xf = xf_lerp(load_xf(rp.cam_xf + i1), load_xf(rp.cam_xf + i1 + 1), a2);
    }
    float tfov = rp.tfov, aperture = rp.aperture, focus = rp.focus;
    if (rp.n_tfov != 0) {
        const float* lens = reinterpret_cast<const float*>(rp.cam_xf + rp.cam_xf_count);
        tfov = lens_lerp(lens, rp.n_tfov, time);
        aperture = lens_lerp(lens + rp.n_tfov, rp.n_aperture, time);
        focus = lens_lerp(lens + rp.n_tfov + rp.n_aperture, rp.n_focus, time);
    }
    float lu, lv;
    square_to_circle((d0.v[1] * 2.0f) - 1.0f, (d0.v[2] * 2.0f) - 1.0f, lu, lv);
    const V3 orig = v3(aperture * lu, aperture * lv, 0.0f);
    const V3 dir = normalize(v3((img_x * tfov) - (orig.x / focus), (img_y * tfov) - (orig.y / focus), 1.0f));
    ps.color = s4_zero();
    ps.atten = s4_splat(1.0f);
    ps.pending = s4_zero();
    ps.csp = 1.0f;
    ps.wl = wl;
    ps.time = time;
    ps.seed = rp.seed ^ 0x40d4682bu;
    ps.bounces = 0u;
    ps.x = x;
    ps.y = y;
    ps.sample = si;
    ps.has_next = false;
    ps.next_o = v3(0.0f, 0.0f, 0.0f);
    ps.next_d = v3(0.0f, 0.0f, 0.0f);
    ps.next_fac = s4_zero();
    o = xf_point(xf, orig);
    d = xf_vector(xf, dir);
}

// Camera::generate_ray and LightPath::new for every slot of the wave.
__global__ __launch_bounds__(KBLOCK) void camera_kernel(const WavePlan* plans, const uint32_t* wave_idx,
                                                        const ReqParams* reqs, ShadeScene sc, Paths P, Queue qa,
                                                        uint32_t* shadow_count) {
    const WavePlan& plan = plans[*wave_idx];
    const uint32_t total = plan.total;
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) {
        *qa.count = total;
        *shadow_count = 0u;
    }
    for (uint32_t i = i0; i < total; i += gridDim.x * blockDim.x) {
        const WaveSeg& sg = find_seg(plan, i);
        const ReqParams rp = reqs[sg.req];
        PathRegs ps;
        V3 o, d;
        camera_path(sc, rp, sg.g0 + (i - sg.slot0), ps, o, d);
        P.req[i] = sg.req;
        P.pix[i] = ps.y * rp.width + ps.x;
        P.sample[i] = ps.sample;
        P.seed[i] = ps.seed;
        P.bounces[i] = 0u;
        P.wl[i] = ps.wl;
        P.time[i] = ps.time;
        P.csp[i] = 1.0f;
        s4_store(P.color, i, ps.color);
        s4_store(P.atten, i, ps.atten);
        s4_store(P.pending, i, ps.pending);
        P.has_next[i] = 0u;
        queue_put(qa, i, i, o, d, ps.time, __builtin_huge_valf());
    }
}

// Shadow-ray results (renderer.rs:670-690): a miss adds the pending light.
__global__ __launch_bounds__(KBLOCK) void shadow_kernel(Queue qs, const uint8_t* kind, Paths P,
                                                        unsigned long long* rays) {
    const uint32_t n = *qs.count;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(rays, (unsigned long long)n);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (kind[i] != PSY_HIT_MISS) continue;
        const uint32_t p = qs.path[i];
        const S4 c = s4_load(P.color, p) + s4_load(P.pending, p);
        s4_store(P.color, p, c);
    }
}

// Sort key of a closest hit: 0 miss, 1 light surface, 2 + closure kind
// (below NUM_BINS for every PSY_SHADER_* kind).
__device__ __forceinline__ uint32_t hit_key(const ShadeScene& sc, const Hits& h, uint32_t q) {
    if (h.kind[q] != PSY_HIT_SURFACE) return 0u;
    const GpuInstance inst = sc.instances[sc.inst_pos[h.inst[q]]];
    if (inst.kind != PSY_OBJECT_MESH) return 1u;
    return 2u + (inst.shader == PSY_NONE ? sc.unassigned.kind : sc.shaders[inst.shader].kind);
}

// Counting sort of the closest hits by key, in two kernels over the same
// block partition: block b takes queue entries [b*chunk, (b+1)*chunk).
// sort_count adds each block's per-key counts to bins[0, NUM_BINS);
// sort_scatter reserves each block's range in every bin with one atomic on
// bins[NUM_BINS, 2 * NUM_BINS) and places the block's entries with LDS
// counters.
__device__ __forceinline__ uint32_t sort_chunk(uint32_t n) {
    return (n + gridDim.x - 1) / gridDim.x;
}

__global__ __launch_bounds__(KBLOCK) void sort_count_kernel(ShadeScene sc, Hits h, const uint32_t* count,
                                                            uint32_t* bins) {
    __shared__ uint32_t local[NUM_BINS];
    if (threadIdx.x < NUM_BINS) local[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = *count;
    const uint32_t chunk = sort_chunk(n);
    const uint32_t q0 = blockIdx.x * chunk, q1 = min(n, q0 + chunk);
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) atomicAdd(&local[hit_key(sc, h, q)], 1u);
    __syncthreads();
    if (threadIdx.x < NUM_BINS && local[threadIdx.x] != 0u) atomicAdd(&bins[threadIdx.x], local[threadIdx.x]);
}

__global__ __launch_bounds__(KBLOCK) void sort_scatter_kernel(ShadeScene sc, Hits h, const uint32_t* count,
                                                              uint32_t* bins, uint32_t* order) {
    __shared__ uint32_t local[NUM_BINS], base[NUM_BINS];
    if (threadIdx.x < NUM_BINS) local[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = *count;
    const uint32_t chunk = sort_chunk(n);
    const uint32_t q0 = blockIdx.x * chunk, q1 = min(n, q0 + chunk);
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) atomicAdd(&local[hit_key(sc, h, q)], 1u);
    __syncthreads();
    if (threadIdx.x < NUM_BINS) {
        uint32_t start = 0;
        for (uint32_t b = 0; b < threadIdx.x; b++) start += bins[b];
        base[threadIdx.x] = local[threadIdx.x] ? start + atomicAdd(&bins[NUM_BINS + threadIdx.x], local[threadIdx.x])
                                               : 0u;
        local[threadIdx.x] = 0u;
    }
    __syncthreads();
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const uint32_t key = hit_key(sc, h, q);
        order[base[key] + atomicAdd(&local[key], 1u)] = q;
    }
}

// Surface intersection data (SurfaceIntersectionData, src/surface/mod.rs:
// 95-105) rebuilt from a hit record as the reference's mesh and light
// intersection code computes it (triangle_mesh.rs:174-310,
// rectangle_light.rs:288-330, sphere_light.rs:283-330).
struct Isect {
    V3 pos, nor, nor_g;
    float pos_err;
    float sample_pdf;
    bool light;
    GpuColor emit;
    uint32_t shader;
};

// The shading helpers below are inlined into every kernel build: a helper
// compiled as a call keeps one register allocation for all callers, so the
// kernels' waves-per-SIMD budgets would not apply to it.
__device__ __forceinline__ Isect make_isect(const ShadeScene& sc, const GpuInstance& inst, float t, float b0,
                                            float b1, float b2, uint32_t prim, V3 orig, V3 dir, float time) {
    Isect is;
    is.light = inst.kind != PSY_OBJECT_MESH;
    is.shader = inst.shader;
    is.sample_pdf = 0.0f;
    if (!is.light) {
        const GpuMesh mesh = sc.meshes[inst.mesh];
        const uint32_t rec = sc.tri_rec[mesh.rec_map + prim];
        const float* f = reinterpret_cast<const float*>(sc.tris + mesh.tri_offset + (uint64_t)rec * mesh.tri_stride);
        V3 v0, v1, v2;
        uint32_t i1;
        float a2;
        if (!lerp_slice_pos(mesh.time_samples, time, i1, a2)) {
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
        // mat_space: the instance transform's inverse, or the identity
        // (triangle_mesh.rs:134-138, 183-189).
        Xf space = xf_identity();
        if (inst.xf_count > 0) {
            space = inst.xf_count == 1 ? load_xf(sc.inv_static + inst.inv_static)
                                       : xf_inverse(instance_xf(sc.xforms, inst, time));
            v0 = xf_point(space, v0);
            v1 = xf_point(space, v1);
            v2 = xf_point(space, v2);
        }
        surface_point(v0, v1, v2, b0, b1, b2, is.pos, is.pos_err);
        is.nor_g = cross(sub(v0, v1), sub(v0, v2));
        is.nor = is.nor_g;
        if (mesh.normal_base != PSY_NONE) {
            // Interpolated shading normal, flipped to the geometric side
            // (triangle_mesh.rs:272-296).
            const uint32_t* ix = sc.nidx + 3ull * ((uint64_t)mesh.nidx_base + prim);
            const uint32_t S = mesh.time_samples;
            auto vnormal = [&](uint32_t vi) {
                const float* n = sc.normals + 3ull * ((uint64_t)mesh.normal_base + (uint64_t)vi * S);
                V3 r;
                if (!lerp_slice_pos(S, time, i1, a2)) {
                    r = v3(n[3 * i1], n[3 * i1 + 1], n[3 * i1 + 2]);
                } else {
                    r = lerp3(v3(n[3 * i1], n[3 * i1 + 1], n[3 * i1 + 2]),
                              v3(n[3 * i1 + 3], n[3 * i1 + 4], n[3 * i1 + 5]), a2);
                }
                return normalize(r);
            };
            const V3 n0 = vnormal(ix[0]), n1 = vnormal(ix[1]), n2 = vnormal(ix[2]);
            const V3 s_nor = xf_normal(space, add(add(mul(n0, b0), mul(n1, b1)), mul(n2, b2)));
            is.nor = dot(s_nor, is.nor_g) >= 0.0f ? s_nor : neg3(s_nor);
        }
        return is;
    }
    const Xf xform = instance_xf(sc.xforms, inst, time);
    if (inst.kind == PSY_OBJECT_RECT_LIGHT) {
        const GpuRectLight L = rect_at(sc.rects[inst.mesh], sc.light_samples, time);
        V3 p[4];
        rect_corners(L, xf_inverse(xform), p);
        const V3 t0 = prim == 0 ? p[0] : p[2], t1 = prim == 0 ? p[1] : p[3], t2 = prim == 0 ? p[2] : p[0];
        surface_point(t0, t1, t2, b0, b1, b2, is.pos, is.pos_err);
        is.nor = cross(sub(t0, t1), sub(t0, t2));
        is.nor_g = is.nor;
        is.sample_pdf = rect_sample_pdf(L, xform, orig, dir, is.pos);
        is.emit = color_scaled(L.color, (float)(1.0 / ((double)L.w * (double)L.h)));
    } else {
        const GpuSphereLight L = sphere_at(sc.spheres[inst.mesh], sc.light_samples, time);
        const V3 lo = xf_point(xform, orig);
        const V3 ld = xf_vector(xform, dir);
        const Xf inv = xf_inverse(xform);
        const V3 unit_pos = normalize(add(lo, mul(ld, t)));
        // `unit_pos * radius * inv_xform` is a Vector times a Transform in the
        // reference: no translation.
        is.pos = xf_vector(inv, mul(unit_pos, L.radius));
        is.pos_err = 0.001f;
        is.nor = xf_normal(inv, unit_pos);
        is.nor_g = is.nor;
        is.sample_pdf = sphere_sample_pdf(L, xform, orig);
        is.emit = color_scaled(L.color, (float)(1.0 / (((4.0 * PI_D) * (double)L.radius) * (double)L.radius)));
    }
    return is;
}

// lerp_slice over a transform list of the light leaves.
__device__ __forceinline__ Xf light_xf_at(const GpuXform* l, uint32_t n, float time) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(n, time, i1, a2)) return load_xf(l + i1);
    return xf_lerp(load_xf(l + i1), load_xf(l + i1 + 1), a2);
}

// LightTree::select (src/accel/light_tree.rs:135-201) in the tree rooted at
// `node`, for shading data already in that tree's space; returns the leaf
// node, with the selection probability and the whittled number.
__device__ __forceinline__ uint32_t light_tree_descend(const ShadeScene& sc, uint32_t node, V3 inc, V3 pos, V3 nor,
                                                       V3 nor_g, const GpuClosure& clo, float time, float& n,
                                                       float& prob) {
    float tot_prob = 1.0f;
    while (sc.lnodes[node].child_count > 0) {
        const GpuLightNode nd = sc.lnodes[node];
        const uint32_t cnt = nd.child_count;
        // The children's probabilities, indexed only by constants (unrolled
        // loops) so they stay in registers.
        float ps[8];
        float total = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; k++) ps[k] = 0.0f;
        for (uint32_t i = 0; i < cnt; i++) {
            const GpuLightNode ch = sc.lnodes[nd.child_first + i];
            // node_prob: energy / bounding-sphere area * closure estimate
            Box b;
            {
                const GpuBox* bs = sc.lboxes + ch.bounds_start;
                uint32_t i1;
                float a2;
                auto ld = [](const GpuBox& g) { return Box{v3(g.lo[0], g.lo[1], g.lo[2]), v3(g.hi[0], g.hi[1], g.hi[2])}; };
                if (!lerp_slice_pos(ch.bounds_len, time, i1, a2)) b = ld(bs[i1]);
                else b = box_lerp(ld(bs[i1]), ld(bs[i1 + 1]), a2);
            }
            const V3 d = sub(box_center(b), pos);
            const float r2 = length2(sub(b.hi, b.lo)) * 0.25f;
            const float inv_surface_area = 1.0f / r2;
            const float approx = closure_estimate_sphere_light(clo, inc, d, r2, nor, nor_g);
            const float p = (ch.energy * inv_surface_area) * approx;
#pragma unroll
            for (int k = 0; k < 8; k++) ps[k] = (uint32_t)k == i ? p : ps[k];
            total += p;
        }
        if (total <= 0.0f) {
            const float p = 1.0f / (float)cnt;
#pragma unroll
            for (int k = 0; k < 8; k++) ps[k] = p;
        } else {
#pragma unroll
            for (int k = 0; k < 8; k++) ps[k] /= total;
        }
        float base = 0.0f;
        bool picked = false;
        uint32_t next = node;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (!picked) {
                const float p = ps[k];
                if ((n <= base + p) || ((uint32_t)k == cnt - 1)) {
                    tot_prob *= p;
                    next = nd.child_first + (uint32_t)k;
                    n = (n - base) / p;
                    picked = true;
                } else {
                    base += p;
                }
            }
        }
        node = next;
    }
    prob = tot_prob;
    return node;
}

// Assembly::sample_lights' selection (src/scene/assembly.rs:46-141): each
// assembly's tree sees the world-space shading data through the lerped
// transforms into that assembly (the identity at the root, which the
// reference applies all the same); a sub-assembly leaf continues in its
// tree with the whittled number.  The reference multiplies the levels'
// probabilities as its recursion returns, innermost first.  Returns the
// light leaf.
__device__ __forceinline__ bool light_tree_select(const ShadeScene& sc, V3 inc, V3 pos, V3 nor, V3 nor_g,
                                                  const GpuClosure& clo, float time, float n, uint32_t& leaf,
                                                  float& prob) {
    if (!sc.has_ltree) return false;
    Xf sel = xf_identity();
    uint32_t node = 0;
    // Each level's probability, indexed only by constants (registers).
    float pl[MAX_LIGHT_LEVELS];
#pragma unroll
    for (int k = 0; k < (int)MAX_LIGHT_LEVELS; k++) pl[k] = 0.0f;
    for (uint32_t level = 0; level < MAX_LIGHT_LEVELS; level++) {
        float pcur;
        node = light_tree_descend(sc, node, xf_vector(sel, inc), xf_point(sel, pos), xf_normal(sel, nor),
                                  xf_normal(sel, nor_g), clo, time, n, pcur);
        const GpuLightLeaf lf = sc.lleaves[sc.lnodes[node].light];
        if (lf.kind == LIGHT_LEAF_LIGHT) {
            float p = pcur;
#pragma unroll
            for (int k = (int)MAX_LIGHT_LEVELS - 1; k >= 0; k--)
                if ((uint32_t)k < level) p = p * pl[k];
            leaf = sc.lnodes[node].light;
            prob = p;
            return true;
        }
#pragma unroll
        for (int k = 0; k < (int)MAX_LIGHT_LEVELS; k++) pl[k] = (uint32_t)k == level ? pcur : pl[k];
        sel = lf.top_count ? light_xf_at(sc.lxforms + lf.top_start, lf.top_count, time) : xf_identity();
        node = lf.target;
    }
    return false;
}

// Scene::sample_lights (src/scene/mod.rs:28-101) and Assembly::sample_lights
// (src/scene/assembly.rs:46-141) for a single-level assembly.
struct LightPick {
    bool found;
    bool distant;
    S4 color;
    V3 dir;  // distant
    V3 point, normal;
    float point_err;
    float pdf, selection_pdf;
};

__device__ __forceinline__ LightPick sample_lights(const ShadeScene& sc, float n, float u, float v, float wl,
                                                   float time, const Isect& is, V3 incoming, const GpuClosure& clo) {
    LightPick out;
    out.found = false;
    float wl_sum = 0.0f;
    for (uint32_t i = 0; i < sc.distant_count; i++) wl_sum = wl_sum + sc.distant[i].energy;
    const float wl_energy = wl_sum <= 0.0f ? 0.0f : 1.0f;
    const float ll_energy = (sc.has_ltree ? sc.lnodes[0].energy : 0.0f) <= 0.0f ? 0.0f : 1.0f;
    const float tot = wl_energy + ll_energy;
    if (tot <= 0.0f) return out;
    const float wl_prob = wl_energy / tot;
    if (n < wl_prob) {
        const float nn = (n / wl_prob) * wl_sum;  // weighted_choice over approximate energies
        float x = 0.0f;
        uint32_t pick = sc.distant_count - 1;
        float w = 0.0f;
        for (uint32_t i = 0; i < sc.distant_count; i++) {
            w = sc.distant[i].energy;
            x += w;
            if (x > nn) {
                pick = i;
                break;
            }
        }
        w = sc.distant[pick].energy;
        const float p = w / wl_sum;
        distant_sample(distant_at(sc.distant[pick], sc.distant_samples, time), u, v, wl, out.color, out.dir, out.pdf);
        out.distant = true;
        out.selection_pdf = p * wl_prob;
        out.found = true;
        return out;
    }
    const float nn = (n - wl_prob) / (1.0f - wl_prob);
    uint32_t li;
    float sel_pdf;
    if (!light_tree_select(sc, incoming, is.pos, is.nor, is.nor_g, clo, time, nn, li, sel_pdf)) return out;
    // The light is sampled under lerp(stack top) * lerp(its own transforms)
    // (assembly.rs:77-95).
    const GpuLightLeaf lf = sc.lleaves[li];
    const GpuInstance inst = sc.instances[lf.target];
    Xf xform = xf_identity();
    if (lf.own_count) {
        xform = light_xf_at(sc.lxforms + lf.own_start, lf.own_count, time);
        if (lf.top_count) xform = xf_then(light_xf_at(sc.lxforms + lf.top_start, lf.top_count, time), xform);
    } else if (lf.top_count) {
        xform = light_xf_at(sc.lxforms + lf.top_start, lf.top_count, time);
    }
    const LightSample s = inst.kind == PSY_OBJECT_RECT_LIGHT
                              ? rect_sample_from_point(rect_at(sc.rects[inst.mesh], sc.light_samples, time), xform,
                                                       is.pos, u, v, wl)
                              : sphere_sample_from_point(sphere_at(sc.spheres[inst.mesh], sc.light_samples, time),
                                                         xform, is.pos, u, v, wl);
    out.found = true;
    out.distant = false;
    out.color = s.color;
    out.point = s.point;
    out.normal = s.normal;
    out.point_err = s.point_err;
    out.pdf = s.pdf;
    out.selection_pdf = sel_pdf * (1.0f - wl_prob);
    return out;
}

// A traced ray's result as the shading step reads it.
struct HitRec {
    uint8_t kind;
    float t, b0, b1, b2;
    uint32_t prim, inst;
};
// What the shading step asks to trace next.
// What one shading step leaves for its caller, so that the path's colour
// and its kept bounce ray need not stay in registers through the step (the
// wavefront kernel reads them from path memory only when used):
//   * `added`: the path's colour gets `add` (emission or background);
//   * `kept`: the path continues along the bounce ray kept from an earlier
//     step (its next_o, next_d), its attenuation times next_fac;
//   * `fresh`: a new bounce ray was drawn into ps.next_o/next_d/next_fac.
struct ShadeOut {
    bool shadow, bounce, added, kept, fresh;
    V3 so, sd, bo, bd;
    float smax;
    S4 add;
};

// LightPath::next for a camera or bounce ray (renderer.rs:459-667): adds
// emission or background, or prepares the light ray (with its pending
// contribution in ps.pending) and the next bounce ray.  DEFER (the wavefront
// kernel) leaves the colour addition and the kept bounce ray to the caller
// (out.added, out.kept), so ps.color and the kept next_* are never read;
// without it (the one-launch kernel, whose path state lives in registers)
// they are applied here.
template <bool DEFER>
__device__ __forceinline__ void shade_path(const ShadeScene& sc, const ReqParams& rp, PathRegs& ps, const HitRec& h,
                                           V3 orig, V3 dir, ShadeOut& out) {
    out.shadow = false;
    out.bounce = false;
    out.added = false;
    out.kept = false;
    out.fresh = false;
    out.smax = 1.0f;
    const float wl = ps.wl;
    const float time = ps.time;
    if (h.kind != PSY_HIT_SURFACE) {
        // Background (renderer.rs:656-666)
        const S4 bg = color_spectral(sc.background, wl);
        if (DEFER) {
            out.add = (bg * ps.atten) / ps.csp;
            out.added = true;
        } else {
            ps.color = ps.color + ((bg * ps.atten) / ps.csp);
        }
        return;
    }
    const GpuInstance inst = sc.instances[sc.inst_pos[h.inst]];
    const Isect is = make_isect(sc, inst, h.t, h.b0, h.b1, h.b2, h.prim, orig, dir, time);
    GpuClosure clo;
    if (is.light) {
        clo.kind = PSY_SHADER_EMIT;
        clo.color = is.emit;
        clo.roughness = 0.0f;
        clo.fresnel = 0.0f;
        clo.ior = 0.0f;
    } else {
        clo = is.shader == PSY_NONE ? sc.unassigned : sc.shaders[is.shader];
    }
    if (clo.kind == PSY_SHADER_EMIT) {
        const S4 c = color_spectral(clo.color, wl);
        S4 add;
        if (ps.bounces == 0u) {
            add = c;
        } else {
            const float mis_pdf = power_heuristic(ps.csp, is.sample_pdf);
            add = (c * ps.atten) / mis_pdf;
        }
        if (DEFER) {
            out.add = add;
            out.added = true;
        } else {
            ps.color = ps.color + add;
        }
        return;
    }
    S4 la = ps.atten / ps.csp;
    // Light ray
    ps.seed += 1u;
    const S4 ls = sample_4d(sc.hdims, sc.hperms, ps.sample, 0, ps.x, ps.y, ps.seed);
    const LightPick li = sample_lights(sc, ls.v[0], ls.v[1], ls.v[2], wl, time, is, dir, clo);
    if (li.found && li.pdf > 0.0f && li.selection_pdf > 0.0f) {
        S4 att;
        float closure_pdf;
        if (li.distant) {
            closure_evaluate(clo, dir, li.dir, is.nor, is.nor_g, wl, att, closure_pdf);
            out.so = robust_ray_origin(is.pos, is.pos_err, normalize(is.nor_g), li.dir);
            out.sd = li.dir;
            out.smax = __builtin_huge_valf();
        } else {
            const V3 ldir = sub(li.point, is.pos);
            closure_evaluate(clo, dir, ldir, is.nor, is.nor_g, wl, att, closure_pdf);
            const V3 offset_pos = robust_ray_origin(is.pos, is.pos_err, normalize(is.nor_g), ldir);
            const V3 offset_end = robust_ray_origin(li.point, li.point_err, normalize(li.normal), neg3(ldir));
            out.so = offset_pos;
            out.sd = sub(offset_end, offset_pos);
            out.smax = 1.0f;
        }
        if (max_element4(att) > 0.0f) {
            const float light_mis_pdf = power_heuristic(li.pdf, closure_pdf);
            ps.pending = ((li.color * att) * la) / (light_mis_pdf * li.selection_pdf);
            out.shadow = true;
        }
    }
    // Bounce ray
    bool do_bounce = false;
    if (ps.bounces < rp.max_bounces) {
        ps.bounces += 1u;
        ps.seed += 1u;
        const S4 bs = sample_4d(sc.hdims, sc.hperms, ps.sample, 0, ps.x, ps.y, ps.seed);
        V3 ndir;
        S4 filter;
        float pdf;
        closure_sample(clo, dir, is.nor, is.nor_g, bs.v[0], bs.v[1], bs.v[2], wl, ndir, filter, pdf);
        if ((pdf > 0.0f) && (max_element4(filter) > 0.0f)) {
            ps.next_fac = filter;
            ps.csp = pdf;
            ps.next_o = robust_ray_origin(is.pos, is.pos_err, normalize(is.nor_g), ndir);
            ps.next_d = ndir;
            ps.has_next = true;
            do_bounce = true;
            out.fresh = true;
        }
    } else {
        ps.has_next = false;
    }
    // A path with a light ray continues after it along next_bounce_ray if it
    // has one, even one kept from an earlier bounce; without a light ray only
    // a fresh bounce continues it.
    if (out.shadow ? ps.has_next : do_bounce) {
        if (!DEFER || do_bounce) {
            out.bo = ps.next_o;
            out.bd = ps.next_d;
            la = la * ps.next_fac;
        } else {
            // along the bounce ray kept from an earlier step: the caller
            // reads it and multiplies the attenuation by its next_fac
            out.kept = true;
        }
        out.bounce = true;
    }
    ps.atten = la;
}

// The shading step over the closest-hit queue (in material order when
// `order` is given); appends the light and bounce rays to the next queues.
// MINW: minimum waves per SIMD the build is compiled for (its register
// budget: 1 and 2 allow 256 VGPRs, 3 about 168, 4 128); every build computes
// the same thing.
template <int MINW>
__global__ __launch_bounds__(KBLOCK) __attribute__((amdgpu_waves_per_eu(MINW))) void shade_kernel(
    ShadeScene sc, const ReqParams* reqs, Paths P, Queue qa, Hits h, const uint32_t* order, Queue qa_out,
    Queue qs_out) {
    const uint32_t n = *qa.count;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(sc.rays, (unsigned long long)n);
    const uint32_t stride = gridDim.x * blockDim.x;
    // Every thread of a block runs the loop the same number of times, so the
    // block-wide appends see all of them.  One atomic per block and queue:
    // per wave, the appends of a launch over a few million rays contend on
    // the two queue counters.
    __shared__ uint32_t app_cnt[2 * (KBLOCK / 64)], app_base[2 * (KBLOCK / 64)];
    const uint32_t n_round = (n + (uint32_t)KBLOCK - 1u) / (uint32_t)KBLOCK * (uint32_t)KBLOCK;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
        ShadeOut so;
        so.shadow = false;
        so.bounce = false;
        float ptime = 0.0f;
        uint32_t p = 0;
        if (i < n) {
            const uint32_t q = order ? order[i] : i;
            p = qa.path[q];
            const ReqParams rp = reqs[P.req[p]];
            PathRegs ps;
            ps.color = s4_zero();    // not read by shade_path (out.add)
            ps.atten = s4_load(P.atten, p);
            ps.pending = s4_zero();  // only written
            ps.csp = P.csp[p];
            ps.wl = P.wl[p];
            ps.time = P.time[p];
            ps.seed = P.seed[p];
            ps.bounces = P.bounces[p];
            ps.x = P.pix[p] % rp.width;
            ps.y = P.pix[p] / rp.width;
            ps.sample = P.sample[p];
            ps.has_next = P.has_next[p] != 0u;
            ps.next_o = v3(0.0f, 0.0f, 0.0f);  // the kept ray stays in path memory (so.kept)
            ps.next_d = v3(0.0f, 0.0f, 0.0f);
            ps.next_fac = s4_zero();
            ptime = ps.time;
            const HitRec hr{h.kind[q], h.t[q], h.b0[q], h.b1[q], h.b2[q], h.prim[q], h.inst[q]};
            shade_path<true>(sc, rp, ps, hr, v3(qa.ox[q], qa.oy[q], qa.oz[q]), v3(qa.dx[q], qa.dy[q], qa.dz[q]), so);
            if (so.added) s4_store(P.color, p, s4_load(P.color, p) + so.add);
            if (so.kept) {
                const float4 a = P.next_o[p], b = P.next_d[p];
                so.bo = v3(a.x, a.y, a.z);
                so.bd = v3(b.x, b.y, b.z);
                ps.atten = ps.atten * s4_load(P.next_fac, p);
            }
            s4_store(P.atten, p, ps.atten);
            if (so.shadow) s4_store(P.pending, p, ps.pending);
            P.csp[p] = ps.csp;
            P.seed[p] = ps.seed;
            P.bounces[p] = ps.bounces;
            P.has_next[p] = ps.has_next ? 1u : 0u;
            if (so.fresh) {
                P.next_o[p] = make_float4(ps.next_o.x, ps.next_o.y, ps.next_o.z, 0.0f);
                P.next_d[p] = make_float4(ps.next_d.x, ps.next_d.y, ps.next_d.z, 0.0f);
                s4_store(P.next_fac, p, ps.next_fac);
            }
        }
        uint32_t at_s, at_b;
        block_append2(qs_out.count, so.shadow, qa_out.count, so.bounce, app_cnt, app_base, at_s, at_b);
        if (so.shadow) queue_put(qs_out, at_s, p, so.so, so.sd, ptime, so.smax);
        if (so.bounce) queue_put(qa_out, at_b, p, so.bo, so.bd, ptime, __builtin_huge_valf());
    }
}

// The plain renderer: one launch per render over the whole frame, one lane
// per path (pixel * spp + sample), each path traced and shaded to its end in
// registers (the reference's per-path logic without wavefronts).  A block
// holds the paths of MEGA_BLOCK / spp whole pixels, or, above MEGA_BLOCK spp,
// one pixel's paths in chunks of MEGA_BLOCK; the paths' XYZ meet in LDS and
// one lane per pixel adds them to the pixel in sample order.
constexpr int MEGA_BLOCK = TRACE_BLOCK;
constexpr int MEGA_LDS = TRACE_LDS_DEPTH;
constexpr int MEGA_SPILL = TRACE_STACK_CAPACITY - TRACE_LDS_DEPTH;

__host__ __device__ inline uint32_t mega_pixels_per_block(uint32_t spp) {
    return spp <= (uint32_t)MEGA_BLOCK ? (uint32_t)MEGA_BLOCK / spp : 1u;
}

template <int MINW>
__global__ __launch_bounds__(MEGA_BLOCK) __attribute__((amdgpu_waves_per_eu(MINW))) void megakernel(
    TraceArgs ta, ShadeScene sc, const ReqParams* reqs) {
    __shared__ uint32_t lds_stack[MEGA_LDS * MEGA_BLOCK];
    __shared__ float lds_xyz[3 * MEGA_BLOCK];
    const ReqParams rp = reqs[0];
    const uint32_t pixels = rp.rw * rp.rh;
    const uint32_t spp = rp.spp;
    const uint32_t lane = threadIdx.x;
    const bool small = spp <= (uint32_t)MEGA_BLOCK;
    const uint32_t per_block = mega_pixels_per_block(spp);
    const uint32_t pix0 = blockIdx.x * per_block;
    const uint32_t chunks = small ? 1u : (spp + MEGA_BLOCK - 1) / MEGA_BLOCK;
    uint32_t* lds_col = lds_stack + lane;
    const float fspp = (float)spp;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;  // the pixel's sum, on its summing lane
    TraceCount tc;
    uint32_t traced = 0;
    for (uint32_t ch = 0; ch < chunks; ch++) {
        const uint32_t pixel = small ? pix0 + lane / spp : pix0;
        const uint32_t si = small ? lane % spp : ch * MEGA_BLOCK + lane;
        const bool active = (!small || lane < per_block * spp) && pixel < pixels && si < spp;
        V3 c = v3(0.0f, 0.0f, 0.0f);
        if (active) {
            PathRegs ps;
            V3 o, d;
            camera_path(sc, rp, (uint64_t)pixel * spp + si, ps, o, d);
            for (;;) {
                const TraceHit th = trace_one<false, MEGA_LDS, MEGA_SPILL, true>(ta, o, d, ps.time, __builtin_huge_valf(),
                                                                           false, lds_col, MEGA_BLOCK, tc);
                const HitRec hr{th.kind, th.t, th.b0, th.b1, th.b2, th.prim, th.inst};
                ShadeOut so;
                shade_path<false>(sc, rp, ps, hr, o, d, so);
                traced += so.shadow ? 2u : 1u;
                if (so.shadow) {
                    const TraceHit sh = trace_one<false, MEGA_LDS, MEGA_SPILL, true>(ta, so.so, so.sd, ps.time, so.smax, true,
                                                                               lds_col, MEGA_BLOCK, tc);
                    if (sh.kind == PSY_HIT_MISS) ps.color = ps.color + ps.pending;
                }
                if (!so.bounce) break;
                o = so.bo;
                d = so.bd;
            }
            c = xyz_from_spectral(ps.color, ps.wl);
        }
        lds_xyz[3 * lane] = c.x;
        lds_xyz[3 * lane + 1] = c.y;
        lds_xyz[3 * lane + 2] = c.z;
        __syncthreads();
        // Film accumulation (renderer.rs:299-310) in sample order.
        if (small) {
            if (lane < per_block && pix0 + lane < pixels) {
                for (uint32_t j = lane * spp; j < (lane + 1) * spp; j++) {
                    fx += lds_xyz[3 * j] / fspp;
                    fy += lds_xyz[3 * j + 1] / fspp;
                    fz += lds_xyz[3 * j + 2] / fspp;
                }
            }
        } else if (lane == 0) {
            const uint32_t m = min((uint32_t)MEGA_BLOCK, spp - ch * MEGA_BLOCK);
            for (uint32_t j = 0; j < m; j++) {
                fx += lds_xyz[3 * j] / fspp;
                fy += lds_xyz[3 * j + 1] / fspp;
                fz += lds_xyz[3 * j + 2] / fspp;
            }
        }
        __syncthreads();
    }
    // One atomic per wave for the ray count.
    for (int off = 32; off > 0; off >>= 1) traced += __shfl_xor(traced, off);
    if ((lane & 63u) == 0u && traced) atomicAdd(sc.rays, (unsigned long long)traced);
    const uint32_t mine = pix0 + lane;
    if (lane >= (small ? per_block : 1u) || mine >= pixels) return;
    float* px = rp.film + 3 * region_pixel(rp, mine);
    px[0] = fx;
    px[1] = fy;
    px[2] = fz;
}

// Path colour to XYZ (XYZ::from_spectral_sample, color.rs:520-526), one
// lane per slot; the result replaces the slot's colour.
__global__ __launch_bounds__(KBLOCK) void xyz_kernel(const WavePlan* plans, const uint32_t* wave_idx, Paths P) {
    const uint32_t total = plans[*wave_idx].total;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const V3 c = xyz_from_spectral(s4_load(P.color, i), P.wl[i]);
        P.color[i] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

// Film accumulation (renderer.rs:299-310): the first slot of each pixel's
// run in the wave adds the run's samples in sample order.
__global__ __launch_bounds__(KBLOCK) void film_kernel(const WavePlan* plans, const uint32_t* wave_idx,
                                                      const ReqParams* reqs, Paths P) {
    const WavePlan& plan = plans[*wave_idx];
    const uint32_t total = plan.total;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const WaveSeg& sg = find_seg(plan, i);
        const ReqParams rp = reqs[sg.req];
        const uint64_t g = sg.g0 + (i - sg.slot0);
        if (i != sg.slot0 && g % rp.spp != 0) continue;
        const uint64_t pixel = region_pixel(rp, g / rp.spp);
        const uint32_t end = sg.slot0 + sg.count;
        float* px = rp.film + 3 * pixel;
        float fx = px[0], fy = px[1], fz = px[2];
        const float fspp = (float)rp.spp;
        const uint32_t run = (uint32_t)std::min<uint64_t>(end - i, rp.spp - g % rp.spp);
        for (uint32_t j = i; j < i + run; j++) {
            const float4 c = P.color[j];
            fx += c.x / fspp;
            fy += c.y / fspp;
            fz += c.z / fspp;
        }
        px[0] = fx;
        px[1] = fy;
        px[2] = fz;
    }
}

// ---- host side -------------------------------------------------------------------

// Faure permutations and the per-dimension tables of the Halton sampler
// (sub_crates/halton/build.rs:101-291), computed once per renderer.
void halton_tables(std::vector<HaltonDim>& dims, std::vector<uint16_t>& perms) {
    const int NUM = 128;
    std::vector<uint32_t> primes;
    for (uint32_t c = 2; (int)primes.size() < NUM; c++) {
        bool prime = true;
        for (uint32_t p : primes)
            if (c % p == 0) {
                prime = false;
                break;
            }
        if (prime) primes.push_back(c);
    }
    // faure[b]: Faure's permutation of the digits of base b.
    std::vector<std::vector<uint32_t>> faure(primes.back() + 1);
    faure[1] = {0};
    faure[2] = {0, 1};
    for (uint32_t b = 3; b <= primes.back(); b++) {
        std::vector<uint32_t>& f = faure[b];
        if (b % 2 == 0) {
            const std::vector<uint32_t>& h = faure[b / 2];
            for (uint32_t x : h) f.push_back(2 * x);
            for (uint32_t x : h) f.push_back(2 * x + 1);
        } else {
            const uint32_t c = (b - 1) / 2;
            for (uint32_t x : faure[b - 1]) f.push_back(x >= c ? x + 1 : x);
            f.insert(f.begin() + c, c);
        }
    }
    dims.assign(NUM, HaltonDim{2, 2, 0, 0});
    perms.clear();
    for (int i = 1; i < NUM; i++) {
        const uint64_t base = primes[i];
        uint64_t digits = 1, pow_base = base;
        while (pow_base * base <= 500) {
            pow_base *= base;
            digits++;
        }
        uint64_t max_power = pow_base;
        while (max_power * pow_base < (1ull << 32)) max_power *= pow_base;
        HaltonDim d{(uint32_t)base, (uint32_t)pow_base, (uint32_t)max_power, (uint32_t)perms.size()};
        for (uint64_t j = 0; j < pow_base; j++) {
            // Digits of j, least significant first, permuted and reversed.
            uint64_t idx = j, r = 0;
            for (uint64_t k = 0; k < digits; k++) {
                r = r * base + faure[base][idx % base];
                idx /= base;
            }
            perms.push_back((uint16_t)r);
        }
        dims[i] = d;
    }
}

// psy_finalize_hits: the hit geometry of each traced ray, rebuilt from its
// record as the reference's intersection code computes it (make_isect) plus
// the local_space it stores: the mesh's mat_space (the inverse of the
// instance's world-to-object transform, triangle_mesh.rs:134-138, 183-189)
// or the light instance's transform (rectangle_light.rs:318-328,
// sphere_light.rs:320-330).
struct FinalizeArgs {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *time;
    const uint4* rec;  // hit records, or NULL for the arrays below
    const uint8_t* kind;
    const float *t, *b0, *b1, *b2;
    const uint32_t *prim, *inst;
    float4* out;       // 6 float4 per ray (psy_hit_geometry)
    uint64_t n;
};

__global__ __launch_bounds__(KBLOCK) void finalize_kernel(ShadeScene sc, FinalizeArgs f) {
    const uint64_t stride = (uint64_t)gridDim.x * KBLOCK;
    for (uint64_t i = (uint64_t)blockIdx.x * KBLOCK + threadIdx.x; i < f.n; i += stride) {
        uint32_t kind, prim, inst_id;
        float t, b0, b1, b2;
        if (f.rec) {
            const uint4 r0 = f.rec[2 * i], r1 = f.rec[2 * i + 1];
            t = __uint_as_float(r0.x); b0 = __uint_as_float(r0.y);
            b1 = __uint_as_float(r0.z); b2 = __uint_as_float(r0.w);
            prim = r1.x; inst_id = r1.y; kind = r1.z;
        } else {
            kind = f.kind[i];
            t = f.t[i]; b0 = f.b0[i]; b1 = f.b1[i]; b2 = f.b2[i];
            prim = f.prim[i]; inst_id = f.inst[i];
        }
        float4* o = f.out + 6 * i;
        if (kind != PSY_HIT_SURFACE) {
            for (int k = 0; k < 6; k++) o[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            continue;
        }
        const float time = f.time[i];
        const V3 orig = v3(f.ox[i], f.oy[i], f.oz[i]);
        const V3 dir = v3(f.dx[i], f.dy[i], f.dz[i]);
        const GpuInstance inst = sc.instances[sc.inst_pos[inst_id]];
        const Isect is = make_isect(sc, inst, t, b0, b1, b2, prim, orig, dir, time);
        Xf ls = xf_identity();
        if (inst.kind != PSY_OBJECT_MESH) ls = instance_xf(sc.xforms, inst, time);
        else if (inst.xf_count > 0)
            ls = inst.xf_count == 1 ? load_xf(sc.inv_static + inst.inv_static)
                                    : xf_inverse(instance_xf(sc.xforms, inst, time));
        o[0] = make_float4(is.pos.x, is.pos.y, is.pos.z, is.pos_err);
        o[1] = make_float4(is.nor.x, is.nor.y, is.nor.z, is.nor_g.x);
        o[2] = make_float4(is.nor_g.y, is.nor_g.z, ls.m[0][0], ls.m[0][1]);
        o[3] = make_float4(ls.m[0][2], ls.m[0][3], ls.m[1][0], ls.m[1][1]);
        o[4] = make_float4(ls.m[1][2], ls.m[1][3], ls.m[2][0], ls.m[2][1]);
        o[5] = make_float4(ls.m[2][2], ls.m[2][3], is.sample_pdf, 0.0f);
    }
}

}  // namespace

namespace psy {
namespace render {

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + KBLOCK - 1) / KBLOCK, 8192)); }

}  // namespace render
}  // namespace psy

namespace {

TraceArgs queue_args(const psy_renderer* r, const Queue& q, const Hits& h, bool occlusion) {
    TraceArgs a = scene_args(r->scene);
    a.ox = q.ox; a.oy = q.oy; a.oz = q.oz;
    a.dx = q.dx; a.dy = q.dy; a.dz = q.dz;
    a.time = q.time; a.max_t = q.max_t;
    a.flags = nullptr;
    a.all_occlusion = occlusion ? 1u : 0u;
    a.n = r->capacity;
    a.n_dev = q.count;
    a.kind = h.kind;
    a.t = h.t; a.b0 = h.b0; a.b1 = h.b1; a.b2 = h.b2;
    a.prim_id = h.prim; a.inst_id = h.inst;
    return a;
}

// Stages of the renderer's profile (psy_render_profile).
enum { ST_CAMERA, ST_TRACE, ST_SHADOW_TRACE, ST_SHADOW, ST_SORT, ST_SHADE, ST_FILM };

// Adds a device queue count to a 64-bit tally (profile mode).
__global__ void accum_count_kernel(const uint32_t* count, unsigned long long* acc) { *acc += *count; }

int pool_event(psy_renderer* r, hipEvent_t& e) {
    if (r->ev_used == r->ev_pool.size()) {
        hipEvent_t n = nullptr;
        PSY_HIP(hipEventCreate(&n));
        r->ev_pool.push_back(n);
    }
    e = r->ev_pool[r->ev_used++];
    return PSY_OK;
}

// Runs `launch` (which enqueues one stage's kernels on the renderer's
// stream); in profile mode between two HIP events kept for the stage.
template <typename F>
int staged(psy_renderer* r, int stage, F&& launch) {
    if (!r->profile) return launch();
    hipEvent_t a = nullptr, b = nullptr;
    int rc = pool_event(r, a);
    if (!rc) rc = pool_event(r, b);
    if (rc) return rc;
    PSY_HIP(hipEventRecord(a, r->stream));
    rc = launch();
    if (rc) return rc;
    PSY_HIP(hipEventRecord(b, r->stream));
    r->marks.push_back(psy_renderer::Mark{stage, a, b});
    return PSY_OK;
}

// Enqueues the steps of one wave: camera, `steps` closest/shadow rounds, the
// last shadow round and the film.  The wave's plan is plans[*wave_idx].
int enqueue_wave(psy_renderer* r, uint32_t steps) {
    const hipStream_t s = r->stream;
    const uint32_t grid = grid_for(r->capacity);
    const bool tally = r->profile || r->counted;
    Queue qa_in = r->qa[0], qa_out = r->qa[1], qs_in = r->qs[0], qs_out = r->qs[1];
    int rc = staged(r, ST_CAMERA, [&]() -> int {
        hipLaunchKernelGGL(camera_kernel, dim3(grid), dim3(KBLOCK), 0, s, r->plans, r->wave_idx, r->reqs, r->sc, r->P,
                           qa_in, qs_in.count);
        PSY_HIP(hipGetLastError());
        return PSY_OK;
    });
    if (rc) return rc;
    psy_trace_tuning tune = r->tuning;
    if (tune.kernel == PSY_KERNEL_SIMPLE) tune.kernel = PSY_KERNEL_PERSISTENT;
    auto trace = [&](const Queue& q, const Hits& h, bool occl, unsigned int* work, int stage) -> int {
        TraceArgs a = queue_args(r, q, h, occl);
        a.counters = r->prof_dev + (occl ? 7 : 2);
        int e2 = staged(r, stage, [&]() -> int {
            hipError_t e = launch_tuned(r->scene, tune, r->cu_count, work, a, r->counted, s);
            if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("trace launch: ") + hipGetErrorString(e));
            return PSY_OK;
        });
        if (e2 || !tally) return e2;
        hipLaunchKernelGGL(accum_count_kernel, dim3(1), dim3(1), 0, s, q.count, r->prof_dev + (occl ? 1 : 0));
        PSY_HIP(hipGetLastError());
        return PSY_OK;
    };
    for (uint32_t step = 0; step <= steps; step++) {
        const bool last = step == steps;
        if (!last) {
            rc = trace(qa_in, r->ha, false, r->work, ST_TRACE);
            if (rc) return rc;
        }
        rc = trace(qs_in, r->hs, true, r->work + 16, ST_SHADOW_TRACE);
        if (!rc) rc = staged(r, ST_SHADOW, [&]() -> int {
            hipLaunchKernelGGL(shadow_kernel, dim3(grid), dim3(KBLOCK), 0, s, qs_in, r->hs.kind, r->P, r->sc.rays);
            PSY_HIP(hipGetLastError());
            return PSY_OK;
        });
        if (rc) return rc;
        if (last) break;
        hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(64), 0, s, qa_out.count, qs_out.count, r->counts + COUNTS_BINS);
        PSY_HIP(hipGetLastError());
        const uint32_t* order = nullptr;
        if (r->sort_materials) {
            rc = staged(r, ST_SORT, [&]() -> int {
                const uint32_t sgrid = std::min<uint32_t>(grid, 1024);
                hipLaunchKernelGGL(sort_count_kernel, dim3(sgrid), dim3(KBLOCK), 0, s, r->sc, r->ha, qa_in.count,
                                   r->counts + COUNTS_BINS);
                PSY_HIP(hipGetLastError());
                hipLaunchKernelGGL(sort_scatter_kernel, dim3(sgrid), dim3(KBLOCK), 0, s, r->sc, r->ha, qa_in.count,
                                   r->counts + COUNTS_BINS, r->order);
                PSY_HIP(hipGetLastError());
                return PSY_OK;
            });
            if (rc) return rc;
            order = r->order;
        }
        auto shade = r->shade_waves >= 4 ? shade_kernel<4>
                     : r->shade_waves == 3 ? shade_kernel<3>
                     : r->shade_waves == 2 ? shade_kernel<2>
                                           : shade_kernel<1>;
        rc = staged(r, ST_SHADE, [&]() -> int {
            hipLaunchKernelGGL(shade, dim3(grid), dim3(KBLOCK), 0, s, r->sc, r->reqs, r->P, qa_in, r->ha, order, qa_out,
                               qs_out);
            PSY_HIP(hipGetLastError());
            return PSY_OK;
        });
        if (rc) return rc;
        std::swap(qa_in, qa_out);
        std::swap(qs_in, qs_out);
    }
    rc = staged(r, ST_FILM, [&]() -> int {
        hipLaunchKernelGGL(xyz_kernel, dim3(grid), dim3(KBLOCK), 0, s, r->plans, r->wave_idx, r->P);
        PSY_HIP(hipGetLastError());
        hipLaunchKernelGGL(film_kernel, dim3(grid), dim3(KBLOCK), 0, s, r->plans, r->wave_idx, r->reqs, r->P);
        PSY_HIP(hipGetLastError());
        return PSY_OK;
    });
    if (rc) return rc;
    hipLaunchKernelGGL(wave_advance_kernel, dim3(1), dim3(1), 0, s, r->wave_idx);
    PSY_HIP(hipGetLastError());
    return PSY_OK;
}

}  // namespace

namespace psy {
namespace render {

void drop_graphs(psy_renderer* r) {
    for (auto& g : r->graphs) (void)hipGraphExecDestroy(g.second);
    r->graphs.clear();
}

int launch_set_wave(psy_renderer* r, uint32_t value) {
    hipLaunchKernelGGL(wave_begin_kernel, dim3(1), dim3(1), 0, r->stream, r->wave_idx, value);
    PSY_HIP(hipGetLastError());
    return PSY_OK;
}

int launch_zero(float* p, uint64_t n, hipStream_t s) {
    hipLaunchKernelGGL(zero_film_kernel, dim3(grid_for(n)), dim3(KBLOCK), 0, s, p, n);
    PSY_HIP(hipGetLastError());
    return PSY_OK;
}

// Camera::new (camera.rs:22-72) and the image plane of Renderer::render_job
// (renderer.rs:214-231).
int make_params(const psy_camera_desc* cam, const psy_render_settings* st, ReqParams& rp,
                std::vector<GpuXform>& cam_slots) {
    if (!cam || !st) return fail(PSY_E_INVALID, "NULL argument");
    if (st->width == 0 || st->height == 0 || st->spp == 0) return fail(PSY_E_INVALID, "empty image or zero spp");
    if (st->width > 65535 || st->height > 65535) return fail(PSY_E_RANGE, "width and height must be below 65536");
    const uint64_t pixels = (uint64_t)st->width * st->height;
    if (pixels * st->spp >= (1ull << 48)) return fail(PSY_E_RANGE, "image too large");
    if (st->max_bounces > 64) return fail(PSY_E_RANGE, "max_bounces must be at most 64");
    if (cam->xform_count == 0 || !cam->xforms) return fail(PSY_E_INVALID, "camera needs at least one transform");
    if (cam->xform_count > CAM_XFORMS) return fail(PSY_E_RANGE, "at most 64 camera transforms");
    if ((cam->fov_count && !cam->fovs) || (cam->aperture_count && !cam->aperture_radii) ||
        (cam->focus_count && !cam->focus_distances))
        return fail(PSY_E_INVALID, "camera lens list is NULL");
    if (cam->fov_count > CAM_LENS_SAMPLES || cam->aperture_count > CAM_LENS_SAMPLES ||
        cam->focus_count > CAM_LENS_SAMPLES)
        return fail(PSY_E_RANGE, "at most 64 time samples per camera lens parameter");
    std::memset(&rp, 0, sizeof(rp));
    rp.width = st->width;
    rp.height = st->height;
    rp.spp = st->spp;
    rp.max_bounces = st->max_bounces;
    rp.seed = st->seed;
    rp.cmpx = 1.0f / (float)st->width;
    rp.cmpy = 1.0f / (float)st->height;
    const float min_x = -1.0f, max_x = 1.0f;
    const float max_y = (float)st->height / (float)st->width, min_y = -max_y;
    rp.x_extent = max_x - min_x;
    rp.y_extent = max_y - min_y;
    // Camera::new (camera.rs:31-64): a zero focus distance anywhere disables
    // the lens; fovs become tan(fov/2) as sin/cos.
    std::vector<float> tfovs, apertures, focuses;
    if (cam->fov_count) tfovs.assign(cam->fovs, cam->fovs + cam->fov_count);
    else tfovs.assign(1, cam->fov);
    for (float& f : tfovs) f = std::sin(f / 2.0f) / std::cos(f / 2.0f);
    if (cam->aperture_count) apertures.assign(cam->aperture_radii, cam->aperture_radii + cam->aperture_count);
    else apertures.assign(1, cam->aperture_radius);
    if (cam->focus_count) focuses.assign(cam->focus_distances, cam->focus_distances + cam->focus_count);
    else focuses.assign(1, cam->focus_distance);
    if (std::find(focuses.begin(), focuses.end(), 0.0f) != focuses.end()) {
        apertures.assign(1, 0.0f);
        focuses.assign(1, 1.0f);
    }
    rp.tfov = tfovs[0];
    rp.aperture = apertures[0];
    rp.focus = focuses[0];
    rp.cam_xf_count = cam->xform_count;
    cam_slots.assign(cam->xform_count, GpuXform{});
    for (uint32_t i = 0; i < cam->xform_count; i++) std::memcpy(cam_slots[i].r, cam->xforms + 12ull * i, 48);
    if (tfovs.size() > 1 || apertures.size() > 1 || focuses.size() > 1) {
        rp.n_tfov = (uint32_t)tfovs.size();
        rp.n_aperture = (uint32_t)apertures.size();
        rp.n_focus = (uint32_t)focuses.size();
        std::vector<float> lens(tfovs);
        lens.insert(lens.end(), apertures.begin(), apertures.end());
        lens.insert(lens.end(), focuses.begin(), focuses.end());
        const size_t extra = (lens.size() * sizeof(float) + sizeof(GpuXform) - 1) / sizeof(GpuXform);
        cam_slots.resize(cam->xform_count + extra, GpuXform{});
        std::memcpy(cam_slots.data() + cam->xform_count, lens.data(), lens.size() * sizeof(float));
    }
    rp.rx = 0;
    rp.ry = 0;
    rp.rw = st->width;
    rp.rh = st->height;
    if (st->crop_width != 0 || st->crop_height != 0) {
        if (st->crop_width == 0 || st->crop_height == 0 || (uint64_t)st->crop_x + st->crop_width > st->width ||
            (uint64_t)st->crop_y + st->crop_height > st->height)
            return fail(PSY_E_RANGE, "crop region outside the image");
        rp.rx = st->crop_x;
        rp.ry = st->crop_y;
        rp.rw = st->crop_width;
        rp.rh = st->crop_height;
    }
    return PSY_OK;
}

int run_wave(psy_renderer* r, uint32_t steps) {
    if (r->profile) {
        // Direct launches, then the wave's stage times.
        r->marks.clear();
        r->ev_used = 0;
        int rc = enqueue_wave(r, steps);
        if (rc) return rc;
        PSY_HIP(hipStreamSynchronize(r->stream));
        for (const psy_renderer::Mark& m : r->marks) {
            float ms = 0.0f;
            PSY_HIP(hipEventElapsedTime(&ms, m.a, m.b));
            r->stage_ms[m.stage] += ms;
        }
        r->marks.clear();
        r->ev_used = 0;
        return PSY_OK;
    }
    if (!r->use_graphs || r->counted) return enqueue_wave(r, steps);
    auto it = r->graphs.find(steps);
    if (it == r->graphs.end()) {
        hipGraph_t g = nullptr;
        PSY_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_wave(r, steps);
        hipError_t e = hipStreamEndCapture(r->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        it = r->graphs.emplace(steps, ex).first;
    }
    PSY_HIP(hipGraphLaunch(it->second, r->stream));
    return PSY_OK;
}

}  // namespace render
}  // namespace psy

namespace {

int color_check(const psy_color& c) {
    if (c.kind > PSY_COLOR_TEMPERATURE || !std::isfinite(c.a) || !std::isfinite(c.b) || !std::isfinite(c.c))
        return fail(PSY_E_INVALID, "bad colour");
    return PSY_OK;
}

}  // namespace

extern "C" {

int psy_renderer_create(const psy_scene* scene_c, const psy_world_desc* world, uint64_t max_paths,
                        psy_renderer** out) {
    if (!scene_c || !world || !out) return fail(PSY_E_INVALID, "NULL argument");
    *out = nullptr;
    psy_scene* scene = const_cast<psy_scene*>(scene_c);
    if (max_paths == 0) max_paths = 1ull << 22;
    if (max_paths > (1ull << 30)) return fail(PSY_E_RANGE, "max_paths must be at most 2^30");
    if (world->shader_count > 0 && !world->shaders) return fail(PSY_E_INVALID, "shaders is NULL");
    if (world->distant_light_count > 0 && !world->distant_lights) return fail(PSY_E_INVALID, "distant_lights is NULL");
    for (const GpuInstance& gi : scene->host.instances)
        if (gi.kind == PSY_OBJECT_MESH && gi.shader != PSY_NONE && gi.shader >= world->shader_count)
            return fail(PSY_E_INVALID, "instance " + std::to_string(gi.id) + " has no valid shader");
    std::vector<GpuClosure> shaders;
    for (uint32_t i = 0; i < world->shader_count; i++) {
        const psy_shader_desc& d = world->shaders[i];
        if (d.kind > PSY_SHADER_DIELECTRIC) return fail(PSY_E_INVALID, "unknown shader kind");
        if (d.kind == PSY_SHADER_DIELECTRIC &&
            !(d.roughness >= 1e-4f && d.roughness <= 1.0f && d.ior > 1.0001f && d.ior <= 10.0f))
            return fail(PSY_E_INVALID, "a dielectric needs roughness in [1e-4, 1] and ior in (1.0001, 10]");
        int rc = color_check(d.color);
        if (rc) return rc;
        shaders.push_back(GpuClosure{d.kind, d.roughness, d.fresnel,
                                     d.kind == PSY_SHADER_DIELECTRIC ? d.ior : 0.0f,
                                     GpuColor{d.color.kind, d.color.a, d.color.b, d.color.c}});
    }
    // Distant lights (DistantDiskLight::new, distant_disk_light.rs:24-36),
    // with their time samples after the records and the mean colour energy
    // precomputed (approximate_energy, distant_disk_light.rs:91-96).
    std::vector<GpuDistantLight> distant;
    std::vector<float> distant_samples;
    for (uint32_t i = 0; i < world->distant_light_count; i++) {
        const psy_distant_light_desc& d = world->distant_lights[i];
        const std::string who = "distant light " + std::to_string(i) + ": ";
        if ((d.radius_count && !d.radii) || (d.direction_count && !d.directions) || (d.color_count && !d.colors))
            return fail(PSY_E_INVALID, who + "sample list is NULL");
        if (d.radius_count > 64 || d.direction_count > 64 || d.color_count > 64)
            return fail(PSY_E_RANGE, who + "too many time samples");
        std::vector<psy_color> cs;
        std::string why;
        if (!colour_list(d.color, d.color_count, d.colors, cs, why)) return fail(PSY_E_INVALID, who + why);
        GpuDistantLight g{d.radius, {d.direction[0], d.direction[1], d.direction[2]},
                          GpuColor{cs[0].kind, cs[0].a, cs[0].b, cs[0].c}, (uint32_t)distant_samples.size(), 0u,
                          light_energy(cs), 0.0f};
        const uint32_t nr = pack_samples(d.radius_count, d.radii, 1, &g.radius, distant_samples);
        const uint32_t nd = pack_samples(d.direction_count, d.directions, 3, g.dir, distant_samples);
        const uint32_t nc = pack_colours(cs, distant_samples);
        g.counts = nr | (nd << 8) | (nc << 16);
        if (g.counts == 0) g.samples = 0;
        distant.push_back(g);
    }
    int rc = color_check(world->background);
    if (rc) return rc;
    if (scene->host.max_stack > (uint32_t)TRACE_STACK_CAPACITY)
        return fail(PSY_E_RANGE, "scene BVHs are deeper than the traversal stack holds");
    rc = ensure_uploaded(scene);
    if (rc) return rc;

    psy_renderer* r = new (std::nothrow) psy_renderer();
    if (!r) return fail(PSY_E_NOMEM, "out of host memory");
    r->scene = scene;
    r->capacity = max_paths;
    auto bail = [&](int code) {
        delete r;
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&r->ev0);
    if (e == hipSuccess) e = hipEventCreate(&r->ev1);
    if (e != hipSuccess) return bail(fail(PSY_E_DEVICE, std::string("renderer setup: ") + hipGetErrorString(e)));
    rc = device_cu_count(r->cu_count);
    if (rc) return bail(rc);

    const uint64_t C = max_paths;
    Paths& P = r->P;
    rc = r->alloc(&P.req, C);
    if (!rc) rc = r->alloc(&P.pix, C);
    if (!rc) rc = r->alloc(&P.sample, C);
    if (!rc) rc = r->alloc(&P.seed, C);
    if (!rc) rc = r->alloc(&P.bounces, C);
    if (!rc) rc = r->alloc(&P.wl, C);
    if (!rc) rc = r->alloc(&P.time, C);
    if (!rc) rc = r->alloc(&P.csp, C);
    if (!rc) rc = r->alloc(&P.color, C);
    if (!rc) rc = r->alloc(&P.atten, C);
    if (!rc) rc = r->alloc(&P.pending, C);
    if (!rc) rc = r->alloc(&P.has_next, C);
    if (!rc) rc = r->alloc(&P.next_o, C);
    if (!rc) rc = r->alloc(&P.next_d, C);
    if (!rc) rc = r->alloc(&P.next_fac, C);
    for (int k = 0; k < 2 && !rc; k++) rc = r->alloc_queue(r->qa[k]);
    for (int k = 0; k < 2 && !rc; k++) rc = r->alloc_queue(r->qs[k]);
    if (!rc) rc = r->alloc(&r->ha.kind, C);
    float** hf[4] = {&r->ha.t, &r->ha.b0, &r->ha.b1, &r->ha.b2};
    for (int k = 0; k < 4 && !rc; k++) rc = r->alloc(hf[k], C);
    if (!rc) rc = r->alloc(&r->ha.prim, C);
    if (!rc) rc = r->alloc(&r->ha.inst, C);
    if (!rc) rc = r->alloc(&r->hs.kind, C);
    float** sf[4] = {&r->hs.t, &r->hs.b0, &r->hs.b1, &r->hs.b2};
    for (int k = 0; k < 4 && !rc; k++) rc = r->alloc(sf[k], C);
    if (!rc) rc = r->alloc(&r->hs.prim, C);
    if (!rc) rc = r->alloc(&r->hs.inst, C);
    if (!rc) rc = r->alloc(&r->order, C);
    if (!rc) rc = r->alloc(&r->counts, COUNTS_WORDS);
    if (!rc) rc = r->alloc(&r->wave_idx, 1);
    if (!rc) rc = r->alloc(&r->work, 64);
    if (!rc) rc = r->alloc(&r->prof_dev, 16);
    if (!rc && hipMemset(r->prof_dev, 0, 16 * sizeof(unsigned long long)) != hipSuccess)
        rc = fail(PSY_E_DEVICE, "profile counters");
    if (!rc) rc = r->alloc(&r->reqs, MAX_SEGS);
    if (!rc) rc = r->alloc(&r->plans, 2);
    if (!rc) rc = r->alloc(&r->cam, (uint64_t)MAX_SEGS * CAM_SLOTS);
    if (rc) return bail(rc);
    r->req_cap = MAX_SEGS;
    r->plan_cap = 2;
    r->cam_cap = CAM_SLOTS;
    r->qa[0].count = r->counts + 0;
    r->qa[1].count = r->counts + 1;
    r->qs[0].count = r->counts + 2;
    r->qs[1].count = r->counts + 3;

    // Scene data for shading.
    ShadeScene& sc = r->sc;
    sc.instances = scene->instances;
    sc.inst_pos = scene->inst_pos;
    sc.meshes = scene->meshes;
    sc.tris = scene->tris;
    sc.tri_rec = scene->tri_rec;
    sc.normals = scene->normals;
    sc.nidx = scene->nidx;
    sc.rays = reinterpret_cast<unsigned long long*>(r->counts + COUNTS_RAYS);
    sc.xforms = scene->xforms;
    sc.inv_static = scene->inv_static;
    sc.rects = scene->rect_lights;
    sc.spheres = scene->sphere_lights;
    sc.light_samples = scene->light_samples;
    sc.lnodes = scene->light_nodes;
    sc.lboxes = scene->light_boxes;
    sc.lleaves = scene->light_leaves;
    sc.lxforms = scene->light_xforms;
    sc.has_ltree = scene->host.light_nodes.empty() ? 0u : 1u;
    sc.shader_count = (uint32_t)shaders.size();
    {
        // Emit(rec709_to_xyz((1, 0, 1))), the reference's colour for mesh
        // surfaces without a shader (tracer.rs:161-164).
        float m[3];
        psy::color::apply(psy::color::matrices().rec709_to_xyz, 1.0f, 0.0f, 1.0f, m);
        sc.unassigned = GpuClosure{PSY_SHADER_EMIT, 0.0f, 0.0f, 0.0f, GpuColor{PSY_COLOR_XYZ, m[0], m[1], m[2]}};
    }
    sc.distant_count = (uint32_t)distant.size();
    sc.background = GpuColor{world->background.kind, world->background.a, world->background.b, world->background.c};
    GpuClosure* dsh = nullptr;
    GpuDistantLight* ddl = nullptr;
    float* dls = nullptr;
    HaltonDim* hd = nullptr;
    uint16_t* hp = nullptr;
    std::vector<HaltonDim> hdims;
    std::vector<uint16_t> hperms;
    halton_tables(hdims, hperms);
    rc = r->alloc(&dsh, shaders.size());
    if (!rc) rc = r->alloc(&ddl, distant.size());
    if (!rc) rc = r->alloc(&dls, distant_samples.size());
    if (!rc) rc = r->alloc(&hd, hdims.size());
    if (!rc) rc = r->alloc(&hp, hperms.size());
    if (rc) return bail(rc);
    e = hipSuccess;
    if (!shaders.empty()) e = hipMemcpy(dsh, shaders.data(), shaders.size() * sizeof(GpuClosure), hipMemcpyHostToDevice);
    if (e == hipSuccess && !distant.empty())
        e = hipMemcpy(ddl, distant.data(), distant.size() * sizeof(GpuDistantLight), hipMemcpyHostToDevice);
    if (e == hipSuccess && !distant_samples.empty())
        e = hipMemcpy(dls, distant_samples.data(), distant_samples.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(hd, hdims.data(), hdims.size() * sizeof(HaltonDim), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(hp, hperms.data(), hperms.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(fail(PSY_E_DEVICE, std::string("renderer upload: ") + hipGetErrorString(e)));
    sc.shaders = dsh;
    sc.distant = ddl;
    sc.distant_samples = dls;
    sc.hdims = hd;
    sc.hperms = hp;
    *out = r;
    return PSY_OK;
}

int psy_renderer_destroy(psy_renderer* r) {
    delete r;
    return PSY_OK;
}

int psy_render(psy_renderer* r, const psy_camera_desc* cam, const psy_render_settings* st, float* xyz, uint32_t mem,
               psy_render_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (!r || !cam || !st || !xyz) return fail(PSY_E_INVALID, "NULL argument");
    if (mem != PSY_MEM_HOST && mem != PSY_MEM_DEVICE) return fail(PSY_E_INVALID, "mem must be PSY_MEM_HOST or PSY_MEM_DEVICE");
    ReqParams rp;
    std::vector<GpuXform> cx;
    int rc = make_params(cam, st, rp, cx);
    if (rc) return rc;
    const uint64_t pixels = (uint64_t)st->width * st->height;
    const uint64_t total = (uint64_t)rp.rw * rp.rh * st->spp;
    rp.cam_xf = r->cam;

    float* film = xyz;
    float* film_dev = nullptr;
    if (mem == PSY_MEM_HOST) {
        rc = r->alloc(&film_dev, pixels * 3);
        if (rc) return rc;
        film = film_dev;
    }
    rp.film = film;

    // Wave plans: consecutive path ranges of `capacity` slots.
    const uint64_t waves = (total + r->capacity - 1) / r->capacity;
    if (waves > r->plan_cap) {
        // Captured graphs hold the plan array's address.
        drop_graphs(r);
        if (r->plans) r->release(r->plans);
        r->plans = nullptr;
        rc = r->alloc(&r->plans, waves);
        if (rc) return rc;
        r->plan_cap = (uint32_t)waves;
    }
    std::vector<WavePlan> plans(waves);
    for (uint64_t w = 0; w < waves; w++) {
        std::memset(&plans[w], 0, sizeof(WavePlan));
        const uint64_t g0 = w * r->capacity;
        plans[w].nseg = 1;
        plans[w].total = (uint32_t)std::min<uint64_t>(r->capacity, total - g0);
        plans[w].seg[0] = WaveSeg{0, 0, plans[w].total, 0, g0};
    }
    const hipStream_t s = r->stream;
    PSY_HIP(hipMemcpyAsync(r->cam, cx.data(), cx.size() * sizeof(GpuXform), hipMemcpyHostToDevice, s));
    PSY_HIP(hipMemcpyAsync(r->reqs, &rp, sizeof(rp), hipMemcpyHostToDevice, s));
    PSY_HIP(hipMemcpyAsync(r->plans, plans.data(), waves * sizeof(WavePlan), hipMemcpyHostToDevice, s));
    rc = launch_zero(film, pixels * 3, s);
    if (!rc) rc = launch_zero(reinterpret_cast<float*>(r->counts + COUNTS_RAYS), 2, s);  // ray counter
    if (!rc) rc = launch_set_wave(r, 0u);
    if (rc) return rc;
    PSY_HIP(hipEventRecord(r->ev0, s));
    if (r->plain) {
        auto mega = r->plain_waves >= 4 ? megakernel<4>
                    : r->plain_waves == 3 ? megakernel<3>
                    : r->plain_waves == 2 ? megakernel<2>
                                          : megakernel<1>;
        const uint64_t per_block = mega_pixels_per_block(rp.spp);
        const uint64_t blocks = ((uint64_t)rp.rw * rp.rh + per_block - 1) / per_block;
        if (blocks > 0x7fffffffull) return fail(PSY_E_RANGE, "image too large for the one-launch renderer");
        hipLaunchKernelGGL(mega, dim3((unsigned)blocks), dim3(MEGA_BLOCK), 0, s, scene_args(r->scene), r->sc,
                           r->reqs);
        PSY_HIP(hipGetLastError());
    } else {
        for (uint64_t w = 0; w < waves && !rc; w++) rc = run_wave(r, st->max_bounces + 1);
        if (r->profile || r->counted) r->prof_paths += total;
    }
    PSY_HIP(hipEventRecord(r->ev1, s));
    if (rc) {
        (void)hipStreamSynchronize(s);
        if (film_dev) r->release(film_dev);
        return rc;
    }
    if (mem == PSY_MEM_HOST) PSY_HIP(hipMemcpyAsync(xyz, film_dev, pixels * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    PSY_HIP(hipStreamSynchronize(s));
    if (film_dev) r->release(film_dev);
    if (stats) {
        float ms = 0.0f;
        PSY_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
        unsigned long long rays = 0;
        PSY_HIP(hipMemcpy(&rays, r->counts + COUNTS_RAYS, sizeof(rays), hipMemcpyDeviceToHost));
        stats->paths = total;
        stats->rays = rays;
        stats->waves = r->plain ? 0 : waves;
        stats->trace_ms = ms;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return PSY_OK;
}

int psy_renderer_set_options(psy_renderer* r, uint32_t flags) {
    if (!r) return fail(PSY_E_INVALID, "renderer is NULL");
    if (flags & ~(PSY_RENDER_NO_GRAPHS | PSY_RENDER_NO_SORT | PSY_RENDER_PLAIN | PSY_RENDER_SHADE_WAVES_MASK |
                  PSY_RENDER_PROFILE | PSY_RENDER_COUNT))
        return fail(PSY_E_INVALID, "unknown option flag");
    const uint32_t sw = (flags & PSY_RENDER_SHADE_WAVES_MASK) >> PSY_RENDER_SHADE_WAVES_SHIFT;
    if (sw > 4) return fail(PSY_E_INVALID, "shading waves per SIMD must be 1 to 4");
    r->plain = (flags & PSY_RENDER_PLAIN) != 0;
    const bool graphs = (flags & PSY_RENDER_NO_GRAPHS) == 0, sort = (flags & PSY_RENDER_NO_SORT) == 0;
    // The register budget applies to the form being selected: the
    // megakernel with PSY_RENDER_PLAIN, shade_kernel otherwise.
    if (sw != 0 && (flags & PSY_RENDER_PLAIN)) {
        r->plain_waves = sw;
    } else if (sw != 0 && sw != r->shade_waves) {
        drop_graphs(r);
        r->shade_waves = sw;
    }
    if (sort != r->sort_materials) {
        for (auto& g : r->graphs) (void)hipGraphExecDestroy(g.second);
        r->graphs.clear();
    }
    r->use_graphs = graphs;
    r->sort_materials = sort;
    r->profile = (flags & PSY_RENDER_PROFILE) != 0;
    r->counted = (flags & PSY_RENDER_COUNT) != 0;
    return PSY_OK;
}

int psy_renderer_get_profile(psy_renderer* r, psy_render_profile* p) {
    if (!r || !p) return fail(PSY_E_INVALID, "NULL argument");
    unsigned long long c[16] = {};
    PSY_HIP(hipStreamSynchronize(r->stream));
    PSY_HIP(hipMemcpy(c, r->prof_dev, sizeof(c), hipMemcpyDeviceToHost));
    PSY_HIP(hipMemset(r->prof_dev, 0, sizeof(c)));
    std::memset(p, 0, sizeof(*p));
    double* ms[7] = {&p->camera_ms, &p->trace_ms, &p->shadow_trace_ms, &p->shadow_ms, &p->sort_ms, &p->shade_ms,
                     &p->film_ms};
    for (int i = 0; i < 7; i++) {
        *ms[i] = r->stage_ms[i];
        r->stage_ms[i] = 0.0;
    }
    p->paths = r->prof_paths;
    r->prof_paths = 0;
    p->closest_rays = c[0];
    p->shadow_rays = c[1];
    p->closest_node_visits = c[2];
    p->closest_tri_tests = c[4];
    p->shadow_node_visits = c[7];
    p->shadow_tri_tests = c[9];
    p->motion_node_visits = c[3] + c[8];
    p->motion_tri_tests = c[5] + c[10];
    p->instance_visits = c[6] + c[11];
    return PSY_OK;
}

int psy_renderer_autotune(psy_renderer* r, const psy_camera_desc* cam, const psy_render_settings* trial,
                          psy_trace_tuning* best_tuning, uint32_t* best_flags) {
    if (!r || !cam || !trial) return fail(PSY_E_INVALID, "NULL argument");
    ReqParams rp;
    std::vector<GpuXform> cx;
    int rc = make_params(cam, trial, rp, cx);
    if (rc) return rc;
    const uint64_t pixels = (uint64_t)trial->width * trial->height;
    float* film = nullptr;
    rc = r->alloc(&film, pixels * 3);
    if (rc) return rc;
    const psy_trace_tuning saved_t = r->tuning;
    const bool saved_sort = r->sort_materials, saved_plain = r->plain;
    const uint32_t saved_sw = r->shade_waves;
    r->plain = false;
    psy_trace_tuning best_t = saved_t;
    bool best_sort = saved_sort;
    uint32_t best_sw = saved_sw;
    float best_ms = 1e30f;
    // Best of two timed renders after a warm-up (which also captures the
    // wave's graph) of the current choice.
    auto time_trial = [&](float& out_ms) {
        drop_graphs(r);
        out_ms = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            psy_render_stats st;
            const int e = psy_render(r, cam, trial, film, PSY_MEM_DEVICE, &st);
            if (e) return e;
            if (rep > 0) out_ms = std::min(out_ms, (float)st.trace_ms);
        }
        return PSY_OK;
    };
    auto consider = [&](float ms) {
        if (ms < best_ms) {
            best_ms = ms;
            best_t = r->tuning;
            best_sort = r->sort_materials;
            best_sw = r->shade_waves;
        }
    };
    // Stage 1: the traversal launch shape, sorting on, current shading build.
    r->sort_materials = true;
    for (uint32_t sw : {4u, 8u}) {
        for (uint32_t w : {16u, 24u, 32u}) {
            for (uint32_t refill : {8u, 16u, 32u}) {
                for (uint32_t lds : {8u, 16u, 24u}) {
                    if (rc) break;
                    r->tuning = psy_trace_tuning{PSY_KERNEL_PERSISTENT, w, refill, lds, sw, 0};
                    float ms;
                    rc = time_trial(ms);
                    if (!rc) consider(ms);
                }
            }
        }
    }
    // Stage 1b: the traversal's descent threshold.
    for (uint32_t lm : {8u, 16u, 32u, 64u}) {
        if (rc) break;
        r->tuning = best_t;
        r->tuning.leaf_min = lm;
        float ms;
        rc = time_trial(ms);
        if (!rc) consider(ms);
    }
    // Stage 2: the shading kernels' register budget.
    for (uint32_t sw = 1; sw <= 4 && !rc; sw++) {
        r->tuning = best_t;
        r->shade_waves = sw;
        float ms;
        rc = time_trial(ms);
        if (!rc) consider(ms);
    }
    // Stage 3: material sorting off.
    if (!rc) {
        r->tuning = best_t;
        r->shade_waves = best_sw;
        r->sort_materials = false;
        float ms;
        rc = time_trial(ms);
        if (!rc) consider(ms);
    }
    // Stage 4: the one-launch form with each of its register budgets, which
    // wins on small closed scenes where path state never needs to leave
    // registers.
    const uint32_t saved_pw = r->plain_waves;
    uint32_t best_pw = saved_pw;
    float best_plain_ms = 1e30f;
    for (uint32_t pw = 1; pw <= 4 && !rc; pw++) {
        r->tuning = best_t;
        r->shade_waves = best_sw;
        r->sort_materials = best_sort;
        r->plain = true;
        r->plain_waves = pw;
        float ms;
        rc = time_trial(ms);
        r->plain = false;
        if (!rc && ms < best_plain_ms) {
            best_plain_ms = ms;
            best_pw = pw;
        }
    }
    const bool best_plain = best_plain_ms < best_ms;
    r->release(film);
    r->tuning = rc ? saved_t : best_t;
    r->sort_materials = rc ? saved_sort : best_sort;
    r->shade_waves = rc ? saved_sw : best_sw;
    r->plain = rc ? saved_plain : best_plain;
    r->plain_waves = rc ? saved_pw : best_pw;
    drop_graphs(r);
    if (rc) return rc;
    if (best_tuning) *best_tuning = best_t;
    if (best_flags)
        *best_flags = (best_sort ? 0u : PSY_RENDER_NO_SORT) | (best_plain ? PSY_RENDER_PLAIN : 0u) |
                      PSY_RENDER_SHADE_WAVES(best_plain ? best_pw : best_sw);
    return PSY_OK;
}

int psy_finalize_hits(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, const psy_hit_soa* hits,
                      psy_hit_geometry* geometry) {
    if (!tr || !geometry) return fail(PSY_E_INVALID, "NULL argument");
    int rc = check_soa(rays, hits);
    if (rc) return rc;
    if (reinterpret_cast<uintptr_t>(geometry) % 16 != 0) return fail(PSY_E_INVALID, "geometry must be 16-byte aligned");
    if (n == 0) return PSY_OK;
    const psy_scene* s = tr->scene;
    ShadeScene sc{};
    sc.instances = s->instances;
    sc.inst_pos = s->inst_pos;
    sc.meshes = s->meshes;
    sc.tris = s->tris;
    sc.tri_rec = s->tri_rec;
    sc.normals = s->normals;
    sc.nidx = s->nidx;
    sc.xforms = s->xforms;
    sc.inv_static = s->inv_static;
    sc.rects = s->rect_lights;
    sc.spheres = s->sphere_lights;
    sc.light_samples = s->light_samples;
    FinalizeArgs f{};
    f.ox = rays->ox; f.oy = rays->oy; f.oz = rays->oz;
    f.dx = rays->dx; f.dy = rays->dy; f.dz = rays->dz;
    f.time = rays->time;
    f.rec = reinterpret_cast<const uint4*>(hits->records);
    f.kind = hits->kind;
    f.t = hits->t; f.b0 = hits->b0; f.b1 = hits->b1; f.b2 = hits->b2;
    f.prim = hits->prim_id; f.inst = hits->inst_id;
    f.out = reinterpret_cast<float4*>(geometry);
    f.n = n;
    hipLaunchKernelGGL(finalize_kernel, dim3(grid_for(n)), dim3(KBLOCK), 0, tr->stream, sc, f);
    PSY_HIP(hipGetLastError());
    PSY_HIP(hipStreamSynchronize(tr->stream));
    return PSY_OK;
}

int psy_renderer_get_tuning(const psy_renderer* r, psy_trace_tuning* t) {
    if (!r || !t) return fail(PSY_E_INVALID, "NULL argument");
    *t = r->tuning;
    return PSY_OK;
}

int psy_renderer_set_tuning(psy_renderer* r, const psy_trace_tuning* t) {
    if (!r || !t) return fail(PSY_E_INVALID, "NULL argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    if (t->kernel == PSY_KERNEL_SIMPLE) return fail(PSY_E_INVALID, "the renderer traces with a persistent kernel");
    r->tuning = *t;
    for (auto& g : r->graphs) (void)hipGraphExecDestroy(g.second);
    r->graphs.clear();
    return PSY_OK;
}

}  // extern "C"
