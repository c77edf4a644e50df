// Internal objects behind the C ABI handles, shared by the tracer (api.hip)
// and the renderer (render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/psy_rt.h"
#include "scene.h"
#include "trace.h"

namespace psy {
namespace api {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

#define PSY_HIP(call)                                                                                         \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return ::psy::api::fail(PSY_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

template <typename T>
int upload(const std::vector<T>& v, T** out) {
    *out = nullptr;
    if (v.empty()) return PSY_OK;
    PSY_HIP(hipMalloc((void**)out, v.size() * sizeof(T)));
    PSY_HIP(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return PSY_OK;
}

}  // namespace api
}  // namespace psy

struct psy_scene {
    psy::PackedScene host;
    psy::GpuNode* nodes = nullptr;
    psy::GpuBox4* motion = nullptr;
    uint8_t* tris = nullptr;
    psy::GpuMesh* meshes = nullptr;
    psy::GpuInstance* instances = nullptr;
    psy::GpuXform* xforms = nullptr;
    psy::GpuXform* inv_static = nullptr;
    psy::GpuRectLight* rect_lights = nullptr;
    psy::GpuSphereLight* sphere_lights = nullptr;
    float* light_samples = nullptr;
    uint32_t* inst_pos = nullptr;
    uint32_t* tri_rec = nullptr;
    psy::GpuLightNode* light_nodes = nullptr;
    psy::GpuBox* light_boxes = nullptr;
    psy::GpuLightLeaf* light_leaves = nullptr;
    psy::GpuXform* light_xforms = nullptr;
    float* normals = nullptr;
    uint32_t* nidx = nullptr;
    uint8_t* order_table = nullptr;
    psy::GpuGroup* groups = nullptr;
    uint64_t device_bytes = 0;
    bool uploaded = false;
    int device = -1;  // device the scene is resident on

    ~psy_scene() {
        void* bufs[] = {nodes, motion, tris, meshes, instances, xforms, inv_static, rect_lights,
                        sphere_lights, light_samples, inst_pos, tri_rec, light_nodes, light_boxes,
                        light_leaves, light_xforms, normals, nidx, order_table, groups};
        for (void* b : bufs) (void)hipFree(b);
    }
};

struct psy_tracer {
    const psy_scene* scene = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t max_batch = 0;
    uint64_t rays_traced = 0;
    bool timed = false;
    // device staging for host-memory calls: 8 ray floats + flags, hit fields
    void* staging = nullptr;
    unsigned long long* counters = nullptr;
    unsigned int* work = nullptr;   // persistent-kernel ray counter
    psy_trace_tuning tuning{PSY_KERNEL_PERSISTENT, 16, 16, 16, 4};
    uint32_t cu_count = 256;

    ~psy_tracer() {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(staging);
        (void)hipFree(counters);
        (void)hipFree(work);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace psy {
namespace api {

// Uploads the packed scene on first use, so scenes can be built and
// inspected on hosts without a GPU.
int ensure_uploaded(psy_scene* s);
// Scene pointers of the traversal kernels' arguments.
TraceArgs scene_args(const psy_scene* s);
// Compute units of the current device.
int device_cu_count(uint32_t& cus);
// Enqueues one traversal launch with the given launch shape.
hipError_t launch_tuned(const psy_scene* scene, const psy_trace_tuning& tuning, uint32_t cu_count, unsigned int* work,
                        const TraceArgs& a, bool counted, hipStream_t stream);
// Validates a launch shape (PSY_OK or an error with message).
int check_tuning(const psy_trace_tuning* t);
// Ray and hit pointers of the traversal kernels' arguments, and their
// validation.
void set_rays(TraceArgs& a, const psy_ray_soa* r);
void set_hits(TraceArgs& a, const psy_hit_soa* h);
int check_soa(const psy_ray_soa* r, const psy_hit_soa* h);

}  // namespace api
}  // namespace psy
