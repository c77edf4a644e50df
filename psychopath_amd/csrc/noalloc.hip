// psy_trace_noalloc: the stream-ordered trace of device-resident rays with no
// tracer object and no allocation, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, hosts that
// forbid hipMalloc after set-up).
//
// The persistent traversal kernel needs one scratch word: the ray counter its
// waves take rays from.  psy_trace keeps it in the tracer's allocation; here
// it lives in `scratch`, a fixed device array of this library.  Each device
// holds its own copy of the array at its own address, looked up once per
// device with hipGetSymbolAddress; host code only ever passes that address.
// Calls on one device share the array, so calls that overlap on one device
// (two streams, or two host threads) race on the counter.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "api_internal.h"

using namespace psy;
using namespace psy::api;

namespace {

// Scratch of psy_trace_noalloc, per device.  Word 0: the persistent
// kernel's ray counter, reset by a kernel at the start of every call.
static __device__ __align__(256) unsigned int scratch[PSY_NOALLOC_SCRATCH_BYTES / sizeof(unsigned int)];

constexpr int MAX_DEVICES = 64;

struct DeviceScratch {
    std::once_flag once;
    hipError_t err = hipSuccess;
    unsigned int* ptr = nullptr;  // device address of `scratch` on this device
    uint32_t cu_count = 0;
};
DeviceScratch g_scratch[MAX_DEVICES];

// The device address of `scratch` and the CU count of device `dev`, looked
// up on the first call on that device (with it current).
int device_scratch(int dev, unsigned int** ptr, uint32_t* cus) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(PSY_E_RANGE, "device index beyond the 64 the library supports");
    DeviceScratch& d = g_scratch[dev];
    std::call_once(d.once, [&d, dev]() {
        void* p = nullptr;
        d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d.err != hipSuccess) return;
        d.ptr = static_cast<unsigned int*>(p);
        int n = 0;
        d.err = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        d.cu_count = n > 0 ? (uint32_t)n : 256u;
    });
    if (d.err != hipSuccess || !d.ptr)
        return fail(PSY_E_DEVICE, std::string("scratch lookup: ") + hipGetErrorString(d.err));
    *ptr = d.ptr;
    *cus = d.cu_count;
    return PSY_OK;
}

}  // namespace

extern "C" int psy_trace_noalloc(const psy_scene* scene, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits,
                                 const psy_trace_tuning* tuning, void* stream) {
    if (!scene) return fail(PSY_E_INVALID, "scene is NULL");
    int rc = check_soa(rays, hits);
    if (rc) return rc;
    psy_trace_tuning t{PSY_KERNEL_PERSISTENT, 16, 16, 16, 4, 0};
    if (tuning) {
        rc = check_tuning(tuning);
        if (rc) return rc;
        t = *tuning;
    }
    if (t.kernel == PSY_KERNEL_SIMPLE)
        return fail(PSY_E_INVALID, "psy_trace_noalloc runs the persistent kernels only");
    if (n > PSY_NOALLOC_MAX_RAYS)
        return fail(PSY_E_RANGE, "psy_trace_noalloc takes at most 2^31 - 1 rays per call (its ray counter is 32-bit)");
    if (!scene->uploaded)
        return fail(PSY_E_STATE, "the scene is not resident on a device: create a tracer for it first");
    if (scene->host.max_stack > (uint32_t)TRACE_STACK_CAPACITY)
        return fail(PSY_E_RANGE, "scene BVHs are deeper than the traversal stack holds");
    int dev = -1;
    PSY_HIP(hipGetDevice(&dev));
    if (dev != scene->device) return fail(PSY_E_STATE, "the scene is resident on another device than the current one");
    unsigned int* work = nullptr;
    uint32_t cus = 0;
    rc = device_scratch(dev, &work, &cus);
    if (rc) return rc;
    if (n == 0) return PSY_OK;
    TraceArgs a = scene_args(scene);
    set_rays(a, rays);
    set_hits(a, hits);
    a.n = n;
    const hipError_t e = launch_tuned(scene, t, cus, work, a, false, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("trace launch: ") + hipGetErrorString(e));
    return PSY_OK;
}
