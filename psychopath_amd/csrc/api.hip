// C ABI (include/psy_rt.h): scene upload, tracers and device helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "api_internal.h"

using namespace psy;
using namespace psy::api;

namespace psy {
namespace api {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int ensure_uploaded(psy_scene* s) {
    if (s->uploaded) return PSY_OK;
    const PackedScene& h = s->host;
    int r = PSY_OK;
    if (!r) r = upload(h.nodes, &s->nodes);
    if (!r) r = upload(h.motion, &s->motion);
    if (!r) r = upload(h.tris, &s->tris);
    if (!r) r = upload(h.meshes, &s->meshes);
    if (!r) r = upload(h.instances, &s->instances);
    if (!r) r = upload(h.xforms, &s->xforms);
    if (!r) r = upload(h.inv_static, &s->inv_static);
    if (!r) r = upload(h.rect_lights, &s->rect_lights);
    if (!r) r = upload(h.sphere_lights, &s->sphere_lights);
    if (!r) r = upload(h.light_samples, &s->light_samples);
    if (!r) r = upload(h.inst_pos, &s->inst_pos);
    if (!r) r = upload(h.tri_rec, &s->tri_rec);
    if (!r) r = upload(h.light_nodes, &s->light_nodes);
    if (!r) r = upload(h.light_boxes, &s->light_boxes);
    if (!r) r = upload(h.light_leaves, &s->light_leaves);
    if (!r) r = upload(h.light_xforms, &s->light_xforms);
    if (!r) r = upload(h.normals, &s->normals);
    if (!r) r = upload(h.nidx, &s->nidx);
    if (!r) r = upload(h.order_table, &s->order_table);
    if (!r) r = upload(h.groups, &s->groups);
    if (r == PSY_OK) {
        PSY_HIP(hipGetDevice(&s->device));
        s->uploaded = true;
    }
    return r;
}

TraceArgs scene_args(const psy_scene* s) {
    TraceArgs a{};
    a.nodes = s->nodes;
    a.motion = s->motion;
    a.tris = s->tris;
    a.meshes = s->meshes;
    a.instances = s->instances;
    a.xforms = s->xforms;
    a.inv_static = s->inv_static;
    a.rect_lights = s->rect_lights;
    a.sphere_lights = s->sphere_lights;
    a.light_samples = s->light_samples;
    a.order_table = s->order_table;
    a.groups = s->host.nested ? s->groups : nullptr;
    a.top_root = s->host.top_root;
    return a;
}

int device_cu_count(uint32_t& cus_out) {
    int dev = 0, cus = 0;
    PSY_HIP(hipGetDevice(&dev));
    PSY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus > 0) cus_out = (uint32_t)cus;
    return PSY_OK;
}

hipError_t launch_tuned(const psy_scene* scene, const psy_trace_tuning& tuning, uint32_t cu_count, unsigned int* work,
                        const TraceArgs& a, bool counted, hipStream_t stream) {
    if (tuning.kernel == PSY_KERNEL_SIMPLE) return launch_trace(a, counted, stream);
    TraceTuning tt{cu_count * tuning.waves_per_cu, tuning.refill_min, tuning.stack_lds, tuning.simd_waves,
                   tuning.leaf_min, tuning.kernel == PSY_KERNEL_PERSISTENT_COOP ? 1u : 0u};
    return launch_trace_persistent(a, counted, scene->host.has_xf || scene->host.has_lights, scene->host.has_motion,
                                   work, tt, stream);
}

int check_tuning(const psy_trace_tuning* t) {
    if (t->kernel != PSY_KERNEL_PERSISTENT && t->kernel != PSY_KERNEL_SIMPLE && t->kernel != PSY_KERNEL_PERSISTENT_COOP)
        return fail(PSY_E_INVALID, "unknown kernel");
    if (t->waves_per_cu < 1 || t->waves_per_cu > 32) return fail(PSY_E_RANGE, "waves_per_cu must be in [1, 32]");
    if (t->refill_min < 1 || t->refill_min > 64) return fail(PSY_E_RANGE, "refill_min must be in [1, 64]");
    if (t->stack_lds != 8 && t->stack_lds != 16 && t->stack_lds != 24)
        return fail(PSY_E_RANGE, "stack_lds must be 8, 16 or 24");
    if (t->simd_waves != 4 && t->simd_waves != 8) return fail(PSY_E_RANGE, "simd_waves must be 4 or 8");
    if (t->leaf_min > 64) return fail(PSY_E_RANGE, "leaf_min must be in [0, 64]");
    return PSY_OK;
}

void set_rays(TraceArgs& a, const psy_ray_soa* r) {
    a.ox = r->ox; a.oy = r->oy; a.oz = r->oz;
    a.dx = r->dx; a.dy = r->dy; a.dz = r->dz;
    a.time = r->time; a.max_t = r->max_t; a.flags = r->flags;
}
void set_hits(TraceArgs& a, const psy_hit_soa* h) {
    a.rec = reinterpret_cast<uint4*>(h->records);
    a.kind = h->kind;
    a.t = h->t; a.b0 = h->b0; a.b1 = h->b1; a.b2 = h->b2;
    a.prim_id = h->prim_id; a.inst_id = h->inst_id;
}
int check_soa(const psy_ray_soa* r, const psy_hit_soa* h) {
    if (!r || !h) return fail(PSY_E_INVALID, "rays or hits is NULL");
    if (!r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz || !r->time || !r->max_t)
        return fail(PSY_E_INVALID, "a ray array is NULL");
    if (h->records) {
        if (reinterpret_cast<uintptr_t>(h->records) % 16 != 0)
            return fail(PSY_E_INVALID, "hit records must be 16-byte aligned");
        return PSY_OK;
    }
    if (!h->kind || !h->t || !h->b0 || !h->b1 || !h->b2 || !h->prim_id || !h->inst_id)
        return fail(PSY_E_INVALID, "a hit array is NULL");
    return PSY_OK;
}

}  // namespace api
}  // namespace psy

namespace {

int enqueue(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits, bool counted) {
    TraceArgs a = scene_args(tr->scene);
    set_rays(a, rays);
    set_hits(a, hits);
    a.n = n;
    a.counters = tr->counters;
    PSY_HIP(hipEventRecord(tr->ev0, tr->stream));
    hipError_t e = launch_tuned(tr->scene, tr->tuning, tr->cu_count, tr->work, a, counted, tr->stream);
    if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("trace launch: ") + hipGetErrorString(e));
    PSY_HIP(hipEventRecord(tr->ev1, tr->stream));
    tr->timed = true;
    tr->rays_traced += n;
    return PSY_OK;
}

// Host-memory path: stage rays in, trace, copy hits out, in chunks of
// max_batch rays.
int trace_host(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits) {
    const uint64_t B = tr->max_batch;
    if (!tr->staging) {
        // 8 f32 + 1 u8 per ray in; out 1 u8 + 4 f32 + 2 u32 per ray, or a
        // 32-byte record (placed after the SoA part, 16-byte aligned).
        PSY_HIP(hipMalloc(&tr->staging, B * (8 * 4 + 1 + 1 + 6 * 4) + B * sizeof(psy_hit_record) + 256));
    }
    uint8_t* base = static_cast<uint8_t*>(tr->staging);
    float* f = reinterpret_cast<float*>(base);
    psy_ray_soa dr{f, f + B, f + 2 * B, f + 3 * B, f + 4 * B, f + 5 * B, f + 6 * B, f + 7 * B, nullptr};
    float* hf = f + 8 * B;
    uint32_t* hu = reinterpret_cast<uint32_t*>(hf + 4 * B);
    uint8_t* flags = reinterpret_cast<uint8_t*>(hu + 2 * B);
    uint8_t* kind = flags + B;
    const uint64_t rec_off = (B * (8 * 4 + 1 + 1 + 6 * 4) + 15) & ~uint64_t(15);
    psy_hit_record* drec = hits->records ? reinterpret_cast<psy_hit_record*>(base + rec_off) : nullptr;
    psy_hit_soa dh{kind, hf, hf + B, hf + 2 * B, hf + 3 * B, hu, hu + B, drec};

    for (uint64_t s = 0; s < n; s += B) {
        uint64_t m = std::min(B, n - s);
        const float* src[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->time, rays->max_t};
        float* dst[8] = {(float*)dr.ox, (float*)dr.oy, (float*)dr.oz, (float*)dr.dx, (float*)dr.dy, (float*)dr.dz,
                         (float*)dr.time, (float*)dr.max_t};
        for (int k = 0; k < 8; k++)
            PSY_HIP(hipMemcpyAsync(dst[k], src[k] + s, m * 4, hipMemcpyHostToDevice, tr->stream));
        dr.flags = nullptr;
        if (rays->flags) {
            PSY_HIP(hipMemcpyAsync(flags, rays->flags + s, m, hipMemcpyHostToDevice, tr->stream));
            dr.flags = flags;
        }
        int rc = enqueue(tr, &dr, m, &dh, false);
        if (rc) return rc;
        if (drec) {
            PSY_HIP(hipMemcpyAsync(hits->records + s, drec, m * sizeof(psy_hit_record), hipMemcpyDeviceToHost,
                                   tr->stream));
            PSY_HIP(hipStreamSynchronize(tr->stream));
            continue;
        }
        PSY_HIP(hipMemcpyAsync(hits->kind + s, dh.kind, m, hipMemcpyDeviceToHost, tr->stream));
        float* hsrc[4] = {dh.t, dh.b0, dh.b1, dh.b2};
        float* hdst[4] = {hits->t, hits->b0, hits->b1, hits->b2};
        for (int k = 0; k < 4; k++)
            PSY_HIP(hipMemcpyAsync(hdst[k] + s, hsrc[k], m * 4, hipMemcpyDeviceToHost, tr->stream));
        PSY_HIP(hipMemcpyAsync(hits->prim_id + s, dh.prim_id, m * 4, hipMemcpyDeviceToHost, tr->stream));
        PSY_HIP(hipMemcpyAsync(hits->inst_id + s, dh.inst_id, m * 4, hipMemcpyDeviceToHost, tr->stream));
        PSY_HIP(hipStreamSynchronize(tr->stream));
    }
    return PSY_OK;
}

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" {

const char* psy_last_error(void) { return g_error.c_str(); }
const char* psy_version(void) { return "psy_rt 0.1 gfx950"; }

}  // extern "C"

namespace {

int scene_create(const psy_scene_desc* desc, const BuildOptions& opt, psy_scene** out) {
    if (!out) return fail(PSY_E_INVALID, "out is NULL");
    *out = nullptr;
    psy_scene* s = new (std::nothrow) psy_scene();
    if (!s) return fail(PSY_E_NOMEM, "out of host memory");
    std::string err;
    int rc;
    try {
        rc = pack_scene(desc, s->host, err, opt);
    } catch (const std::bad_alloc&) {
        rc = PSY_E_NOMEM;
        err = "out of host memory while building the scene";
    }
    if (rc != PSY_OK) {
        delete s;
        return fail(rc, err);
    }
    const PackedScene& h = s->host;
    s->device_bytes = h.nodes.size() * sizeof(GpuNode) + h.motion.size() * sizeof(GpuBox4) + h.tris.size() +
                      h.meshes.size() * sizeof(GpuMesh) + h.instances.size() * sizeof(GpuInstance) +
                      (h.xforms.size() + h.inv_static.size()) * sizeof(GpuXform) +
                      h.rect_lights.size() * sizeof(GpuRectLight) + h.sphere_lights.size() * sizeof(GpuSphereLight) +
                      h.light_samples.size() * sizeof(float) +
                      (h.inst_pos.size() + h.tri_rec.size()) * 4 + h.light_leaves.size() * sizeof(GpuLightLeaf) +
                      h.light_xforms.size() * sizeof(GpuXform) +
                      h.light_nodes.size() * sizeof(GpuLightNode) + h.light_boxes.size() * sizeof(GpuBox) +
                      (h.normals.size() + h.nidx.size()) * 4 + h.order_table.size() +
                      h.groups.size() * sizeof(GpuGroup);
    *out = s;
    return PSY_OK;
}

}  // namespace

extern "C" {

int psy_scene_create(const psy_scene_desc* desc, psy_scene** out) { return scene_create(desc, BuildOptions{}, out); }

int psy_scene_create_on_device(const psy_scene_desc* desc, uint32_t min_tris, psy_scene** out) {
    BuildOptions opt;
    opt.device = true;
    opt.device_min_tris = min_tris;
    return scene_create(desc, opt, out);
}

int psy_scene_destroy(psy_scene* scene) {
    delete scene;
    return PSY_OK;
}

int psy_scene_get_info(const psy_scene* s, psy_scene_info* info) {
    if (!s || !info) return fail(PSY_E_INVALID, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    info->node_count = s->host.nodes.size();
    info->leaf_count = s->host.leaf_count;
    info->motion_bounds = s->host.motion.size();
    info->tri_count = s->host.tri_count;
    info->device_bytes = s->device_bytes;
    info->node_bytes = sizeof(GpuNode);
    info->motion_bounds_bytes = sizeof(GpuBox4);
    info->tri_bytes = sizeof(GpuTriStatic);
    info->max_depth = s->host.max_depth;
    info->build_seconds = s->host.build_seconds;
    info->device_built_meshes = s->host.device_built;
    info->device_build_ms = s->host.device_build_ms;
    info->device_light_trees = s->host.device_light_trees;
    return PSY_OK;
}

int64_t psy_scene_dump_bvh(const psy_scene* s, int32_t mesh, uint32_t* out, uint64_t cap, uint32_t* order,
                           uint64_t ocap) {
    if (!s) return fail(PSY_E_INVALID, "scene is NULL");
    const auto& bvhs = s->host.bvhs;
    int64_t nmesh = (int64_t)bvhs.size() - 1;
    if (mesh < -1 || mesh >= nmesh) return fail(PSY_E_INVALID, "mesh index out of range");
    const HostBvh4& b = mesh < 0 ? bvhs.back() : bvhs[mesh];
    std::vector<uint32_t> v;
    dump_bvh4(b, v);
    if (out && v.size() <= cap) std::memcpy(out, v.data(), v.size() * 4);
    if (order) std::memcpy(order, b.order.data(), std::min<uint64_t>(ocap, b.order.size()) * 4);
    return (int64_t)v.size();
}

int psy_tracer_create(const psy_scene* scene, uint64_t max_batch, psy_tracer** out) {
    if (!scene || !out) return fail(PSY_E_INVALID, "NULL argument");
    *out = nullptr;
    if (max_batch == 0 || max_batch > (1ull << 31)) return fail(PSY_E_RANGE, "max_batch must be in [1, 2^31]");
    if (scene->host.max_stack > (uint32_t)TRACE_STACK_CAPACITY)
        return fail(PSY_E_RANGE, "scene BVHs are deeper than the traversal stack holds");
    int up = ensure_uploaded(const_cast<psy_scene*>(scene));
    if (up) return up;
    psy_tracer* t = new (std::nothrow) psy_tracer();
    if (!t) return fail(PSY_E_NOMEM, "out of host memory");
    t->scene = scene;
    t->max_batch = max_batch;
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&t->ev0);
    if (e == hipSuccess) e = hipEventCreate(&t->ev1);
    if (e == hipSuccess) e = hipMalloc((void**)&t->counters, 8 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&t->work, 64);
    if (e != hipSuccess) {
        delete t;
        return fail(PSY_E_DEVICE, std::string("tracer setup: ") + hipGetErrorString(e));
    }
    int rc = device_cu_count(t->cu_count);
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return PSY_OK;
}

int psy_tracer_destroy(psy_tracer* tracer) {
    delete tracer;
    return PSY_OK;
}

int psy_trace(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits, uint32_t mem) {
    if (!tr) return fail(PSY_E_INVALID, "tracer is NULL");
    int rc = check_soa(rays, hits);
    if (rc) return rc;
    if (n == 0) return PSY_OK;
    if (mem == PSY_MEM_HOST) return trace_host(tr, rays, n, hits);
    if (mem != PSY_MEM_DEVICE) return fail(PSY_E_INVALID, "mem must be PSY_MEM_HOST or PSY_MEM_DEVICE");
    rc = enqueue(tr, rays, n, hits, false);
    if (rc) return rc;
    PSY_HIP(hipStreamSynchronize(tr->stream));
    return PSY_OK;
}

int psy_trace_async(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits) {
    if (!tr) return fail(PSY_E_INVALID, "tracer is NULL");
    int rc = check_soa(rays, hits);
    if (rc) return rc;
    if (n == 0) return PSY_OK;
    return enqueue(tr, rays, n, hits, false);
}

int psy_tracer_get_tuning(const psy_tracer* tr, psy_trace_tuning* t) {
    if (!tr || !t) return fail(PSY_E_INVALID, "NULL argument");
    *t = tr->tuning;
    return PSY_OK;
}

int psy_tracer_set_tuning(psy_tracer* tr, const psy_trace_tuning* t) {
    if (!tr || !t) return fail(PSY_E_INVALID, "NULL argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    tr->tuning = *t;
    return PSY_OK;
}

int psy_tracer_autotune(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits,
                        psy_trace_tuning* best) {
    if (!tr) return fail(PSY_E_INVALID, "tracer is NULL");
    int rc = check_soa(rays, hits);
    if (rc) return rc;
    if (n == 0) return fail(PSY_E_INVALID, "autotune needs rays");
    std::vector<psy_trace_tuning> cands;
    cands.push_back({PSY_KERNEL_SIMPLE, 1, 1, 8, 4, 0});
    const uint32_t wpc[] = {8, 12, 16, 24, 32};
    const uint32_t refill[] = {1, 8, 16, 32};
    const uint32_t depth[] = {8, 16, 24};
    for (uint32_t w : wpc)
        for (uint32_t r : refill)
            for (uint32_t dd : depth)
                for (uint32_t occ : {4u, 8u}) cands.push_back({PSY_KERNEL_PERSISTENT, w, r, dd, occ, 0});
    // The cooperative node fetch (static scenes only; on scenes with time
    // samples it runs as the plain persistent kernel).  Its 7 KB of node
    // staging per wave caps it near 16 resident waves per CU.
    if (!tr->scene->host.has_motion)
        for (uint32_t w : {8u, 12u, 16u})
            for (uint32_t r : {8u, 16u, 32u})
                for (uint32_t dd : {8u, 16u}) cands.push_back({PSY_KERNEL_PERSISTENT_COOP, w, r, dd, 4u, 0});
    const psy_trace_tuning saved = tr->tuning;
    float best_ms = 1e30f;
    psy_trace_tuning best_t = saved;
    uint64_t runs = 0;
    auto time_all = [&]() -> int {
        for (const psy_trace_tuning& c : cands) {
            tr->tuning = c;
            float ms_min = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                const int e = enqueue(tr, rays, n, hits, false);
                if (e) {
                    tr->tuning = saved;
                    return e;
                }
                float ms = 0;
                PSY_HIP(hipEventSynchronize(tr->ev1));
                PSY_HIP(hipEventElapsedTime(&ms, tr->ev0, tr->ev1));
                if (rep > 0) ms_min = std::min(ms_min, ms);
            }
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best_t = c;
            }
        }
        runs += 3 * cands.size();
        return PSY_OK;
    };
    // Stage 1: kernel, waves, refill, LDS depth and register budget with
    // per-lane descents; stage 2: the in-loop descents and their threshold
    // around the best.
    rc = time_all();
    if (rc) return rc;
    if (best_t.kernel != PSY_KERNEL_SIMPLE) {
        const psy_trace_tuning s1 = best_t;
        cands.clear();
        const uint32_t k = s1.kernel;
        for (uint32_t lm : {8u, 16u, 24u, 32u, 48u, 64u})
            for (uint32_t r : {8u, 16u, 32u})
                for (uint32_t w : (k == PSY_KERNEL_PERSISTENT_COOP ? std::vector<uint32_t>{8u, 12u, 16u}
                                                                   : std::vector<uint32_t>{16u, 24u, 32u}))
                    cands.push_back({k, w, r, s1.stack_lds, s1.simd_waves, lm});
        rc = time_all();
        if (rc) return rc;
    }
    tr->rays_traced -= runs * n;
    tr->tuning = best_t;
    if (best) *best = best_t;
    return PSY_OK;
}

int psy_tracer_sync(psy_tracer* tr) {
    if (!tr) return fail(PSY_E_INVALID, "tracer is NULL");
    PSY_HIP(hipStreamSynchronize(tr->stream));
    return PSY_OK;
}

int psy_tracer_last_kernel_ms(psy_tracer* tr, float* ms) {
    if (!tr || !ms) return fail(PSY_E_INVALID, "NULL argument");
    if (!tr->timed) return fail(PSY_E_STATE, "no kernel has run on this tracer");
    PSY_HIP(hipEventSynchronize(tr->ev1));
    PSY_HIP(hipEventElapsedTime(ms, tr->ev0, tr->ev1));
    return PSY_OK;
}

int psy_trace_counted(psy_tracer* tr, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits,
                      psy_trace_counters* c) {
    if (!tr || !c) return fail(PSY_E_INVALID, "NULL argument");
    int rc = check_soa(rays, hits);
    if (rc) return rc;
    std::memset(c, 0, sizeof(*c));
    if (n == 0) return PSY_OK;
    unsigned long long zero[8] = {};
    PSY_HIP(hipMemcpyAsync(tr->counters, zero, sizeof(zero), hipMemcpyHostToDevice, tr->stream));
    rc = enqueue(tr, rays, n, hits, true);
    if (rc) return rc;
    unsigned long long h[8];
    PSY_HIP(hipMemcpyAsync(h, tr->counters, sizeof(h), hipMemcpyDeviceToHost, tr->stream));
    PSY_HIP(hipStreamSynchronize(tr->stream));
    c->node_visits = h[0];
    c->motion_node_visits = h[1];
    c->tri_tests = h[2];
    c->motion_tri_tests = h[3];
    c->instance_visits = h[4];
    c->rays = n;
    return PSY_OK;
}

uint64_t psy_tracer_rays_traced(const psy_tracer* tr) { return tr ? tr->rays_traced : 0; }

int psy_device_alloc(uint64_t bytes, void** ptr) {
    if (!ptr) return fail(PSY_E_INVALID, "ptr is NULL");
    PSY_HIP(hipMalloc(ptr, bytes));
    return PSY_OK;
}
int psy_device_free(void* ptr) {
    PSY_HIP(hipFree(ptr));
    return PSY_OK;
}
int psy_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
    PSY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PSY_OK;
}
int psy_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
    PSY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PSY_OK;
}
int psy_device_count(int* count) {
    if (!count) return fail(PSY_E_INVALID, "count is NULL");
    PSY_HIP(hipGetDeviceCount(count));
    return PSY_OK;
}

int psy_device_copy_kernel(void* dst, const void* src, uint64_t bytes, float* ms) {
    if (!dst || !src || bytes % 16 != 0) return fail(PSY_E_INVALID, "copy needs non-NULL pointers and bytes % 16 == 0");
    hipEvent_t e0, e1;
    PSY_HIP(hipEventCreate(&e0));
    PSY_HIP(hipEventCreate(&e1));
    uint64_t n = bytes / 16;
    PSY_HIP(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, nullptr, static_cast<const float4*>(src),
                       static_cast<float4*>(dst), n);
    PSY_HIP(hipGetLastError());
    PSY_HIP(hipEventRecord(e1, nullptr));
    PSY_HIP(hipEventSynchronize(e1));
    float t = 0;
    PSY_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return PSY_OK;
}

}  // extern "C"
