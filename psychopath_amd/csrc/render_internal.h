// Internal state of the wavefront renderer, shared by the single-render
// path (render.hip) and the request server (serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/psy_render.h"
#include "api_internal.h"
#include "render_device.h"
#include "trace_device.h"

namespace psy {
namespace render {

using rd::HaltonDim;

constexpr int KBLOCK = 256;
constexpr uint32_t MAX_SEGS = 64;  // requests interleaved in one wave
constexpr uint32_t NUM_BINS = 6;   // miss, light, Lambert, GGX, emit, dielectric
// Layout of the renderer's counter words: [0, 4) queue counts, then the sort
// bins (2 * NUM_BINS), then the rays-traced u64 at an even word.
constexpr uint32_t COUNTS_BINS = 4;
constexpr uint32_t COUNTS_RAYS = 16;
constexpr uint32_t COUNTS_WORDS = COUNTS_RAYS + 2;
static_assert(COUNTS_BINS + 2 * NUM_BINS <= COUNTS_RAYS && COUNTS_RAYS % 2 == 0, "counter words");

// A request's camera on the device: its transforms, then (when any lens
// parameter has several time samples) the lens lists as floats, linear
// fovs first, then aperture radii, then focus distances.
constexpr uint32_t CAM_XFORMS = 64;       // transforms per camera
constexpr uint32_t CAM_LENS_SAMPLES = 64;  // samples per lens list
constexpr uint32_t CAM_SLOTS =            // GpuXform-sized slots per camera
    CAM_XFORMS + (3 * CAM_LENS_SAMPLES * 4 + sizeof(GpuXform) - 1) / sizeof(GpuXform);

// One request's render parameters (device copy).
struct ReqParams {
    uint32_t width, height, spp, max_bounces, seed;
    float cmpx, cmpy, x_extent, y_extent, tfov, aperture, focus;
    uint32_t cam_xf_count;
    // Rendered region (the whole image unless cropped): path g belongs to
    // region pixel g / spp, row-major in the region.
    uint32_t rx, ry, rw, rh;
    // Lengths of the lens lists after the transforms; all 0 when every lens
    // parameter has one sample (then tfov, aperture and focus hold them).
    uint32_t n_tfov, n_aperture, n_focus;
    const GpuXform* cam_xf;
    float* film;
};

// Image pixel index of region pixel `p`.
__host__ __device__ inline uint64_t region_pixel(const ReqParams& rp, uint64_t p) {
    return (uint64_t)(rp.ry + (uint32_t)(p / rp.rw)) * rp.width + rp.rx + (uint32_t)(p % rp.rw);
}

// A run of consecutive path slots of a wave taken by one request: slots
// [slot0, slot0 + count) carry the request's global paths g0, g0 + 1, ...
// (path g = pixel * spp + sample).
struct WaveSeg {
    uint32_t req;
    uint32_t slot0;
    uint32_t count;
    uint32_t pad;
    uint64_t g0;
};
struct WavePlan {
    uint32_t nseg;
    uint32_t total;
    uint32_t pad[2];
    WaveSeg seg[MAX_SEGS];
};

// Scene data the shading kernels read.
struct ShadeScene {
    const GpuInstance* instances;
    const uint32_t* inst_pos;
    const GpuMesh* meshes;
    const uint8_t* tris;
    const uint32_t* tri_rec;
    const float* normals;
    const uint32_t* nidx;
    const GpuXform* xforms;
    const GpuXform* inv_static;
    const GpuRectLight* rects;
    const GpuSphereLight* spheres;
    const float* light_samples;  // time samples of the lights (layout.h)
    const GpuLightNode* lnodes;
    const GpuBox* lboxes;
    const GpuLightLeaf* lleaves;
    const GpuXform* lxforms;
    uint32_t has_ltree;
    uint32_t shader_count;
    const GpuClosure* shaders;
    GpuClosure unassigned;  // mesh surfaces without a shader (tracer.rs:161-164)
    const GpuDistantLight* distant;
    const float* distant_samples;  // time samples of the distant lights
    uint32_t distant_count;
    uint32_t pad;
    GpuColor background;
    const HaltonDim* hdims;
    const uint16_t* hperms;
    unsigned long long* rays;  // rays traced (RenderStats::ray_count), summed by the kernels
};

// Path state, one slot per path of the wave.
struct Paths {
    uint32_t* req;
    uint32_t* pix;
    uint32_t* sample;
    uint32_t* seed;
    uint32_t* bounces;
    float* wl;
    float* time;
    float* csp;  // closure_sample_pdf
    float4* color;
    float4* atten;
    float4* pending;
    uint32_t* has_next;  // a bounce ray is kept (see PathRegs in render.hip)
    float4* next_o;
    float4* next_d;
    float4* next_fac;
};

// A ray queue: the path slot of each ray and the rays as SoA.
struct Queue {
    uint32_t* path;
    float *ox, *oy, *oz, *dx, *dy, *dz, *time, *max_t;
    uint32_t* count;
};
struct Hits {
    uint8_t* kind;
    float *t, *b0, *b1, *b2;
    uint32_t *prim, *inst;
};

struct DevBuf {
    void* p = nullptr;
    uint64_t bytes = 0;
    ~DevBuf() { (void)hipFree(p); }
};

}  // namespace render
}  // namespace psy

struct psy_renderer {
    using ShadeScene = psy::render::ShadeScene;
    using Paths = psy::render::Paths;
    using Queue = psy::render::Queue;
    using Hits = psy::render::Hits;
    using DevBuf = psy::render::DevBuf;
    using WavePlan = psy::render::WavePlan;
    using ReqParams = psy::render::ReqParams;

    psy_scene* scene = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t capacity = 0;
    uint32_t cu_count = 256;
    psy_trace_tuning tuning{PSY_KERNEL_PERSISTENT, 16, 16, 16, 4};
    bool use_graphs = true;
    bool sort_materials = true;
    bool plain = false;  // one megakernel launch per render
    uint32_t shade_waves = 2;  // minimum waves per SIMD of shade_kernel's build (1-4)
    uint32_t plain_waves = 2;  // the same for the one-launch megakernel
    // Measurement (PSY_RENDER_PROFILE / PSY_RENDER_COUNT): events around
    // each kernel of the current wave (stage, start, end), a pool reused
    // wave after wave, and the tallies.  prof_dev: [0] closest rays,
    // [1] shadow rays, [2..7) closest-ray and [7..12) shadow-ray traversal
    // counters (nodes, motion nodes, triangles, motion triangles, instances).
    bool profile = false;
    bool counted = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Mark {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    double stage_ms[7] = {};
    uint64_t prof_paths = 0;
    unsigned long long* prof_dev = nullptr;

    std::vector<DevBuf*> bufs;
    ShadeScene sc{};
    Paths P{};
    Queue qa[2]{}, qs[2]{};
    Hits ha{}, hs{};
    uint32_t* order = nullptr;
    uint32_t* counts = nullptr;  // COUNTS_WORDS words: queue counts, sort bins, rays traced (u64)
    uint32_t* wave_idx = nullptr;
    unsigned int* work = nullptr;
    WavePlan* plans = nullptr;   // device plan array
    uint32_t plan_cap = 0;
    ReqParams* reqs = nullptr;
    uint32_t req_cap = 0;
    psy::GpuXform* cam = nullptr;
    uint32_t cam_cap = 0;

    // Captured wave graphs, keyed by the number of bounce steps.
    std::map<uint32_t, hipGraphExec_t> graphs;

    ~psy_renderer() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        for (DevBuf* b : bufs) delete b;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    uint64_t device_bytes() const {
        uint64_t n = 0;
        for (const DevBuf* b : bufs) n += b->bytes;
        return n;
    }

    template <typename T>
    int alloc(T** out, uint64_t count) {
        DevBuf* b = new (std::nothrow) DevBuf();
        if (!b) return psy::api::fail(PSY_E_NOMEM, "out of host memory");
        b->bytes = std::max<uint64_t>(1, count) * sizeof(T);
        hipError_t e = hipMalloc(&b->p, b->bytes);
        if (e != hipSuccess) {
            delete b;
            return psy::api::fail(PSY_E_NOMEM, std::string("device allocation: ") + hipGetErrorString(e));
        }
        bufs.push_back(b);
        *out = static_cast<T*>(b->p);
        return PSY_OK;
    }
    void release(void* p) {
        for (size_t i = 0; i < bufs.size(); i++)
            if (bufs[i]->p == p) {
                delete bufs[i];
                bufs.erase(bufs.begin() + i);
                return;
            }
    }

    int alloc_queue(Queue& q) {
        int r = alloc(&q.path, capacity);
        float** f[8] = {&q.ox, &q.oy, &q.oz, &q.dx, &q.dy, &q.dz, &q.time, &q.max_t};
        for (int k = 0; k < 8 && !r; k++) r = alloc(f[k], capacity);
        return r;
    }
};


namespace psy {
namespace render {

// Grid of the queue kernels for `n` items (grid-stride loops).
uint32_t grid_for(uint64_t n);
// Enqueues one wave (camera, bounce steps, film) on the renderer's stream,
// replaying or capturing its hipGraph; the wave's plan is
// plans[*wave_idx], and *wave_idx is incremented at the end.
int run_wave(psy_renderer* r, uint32_t steps);
// Kernels the server launches around waves.
int launch_set_wave(psy_renderer* r, uint32_t value);
int launch_zero(float* p, uint64_t n, hipStream_t s);
// Checks the camera and settings and fills `rp` (except cam_xf and film)
// and `cam_slots`, the camera's device image (at most CAM_SLOTS entries).
int make_params(const psy_camera_desc* cam, const psy_render_settings* st, ReqParams& rp,
                std::vector<GpuXform>& cam_slots);
// Drops captured graphs (after a buffer they name was replaced).
void drop_graphs(psy_renderer* r);

}  // namespace render
}  // namespace psy
