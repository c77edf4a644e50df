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
#include "trace_device.h"
#include "trace_one.h"

namespace psy {

namespace {

constexpr int BLOCK = TRACE_BLOCK;
constexpr int LDS_DEPTH = TRACE_LDS_DEPTH;
constexpr int SPILL_DEPTH = TRACE_STACK_CAPACITY - TRACE_LDS_DEPTH;

template <bool COUNT, bool NESTED>
__global__ __launch_bounds__(BLOCK) void trace_kernel(TraceArgs a) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * BLOCK];
    const uint64_t ray = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (ray >= a.n || (a.n_dev != nullptr && ray >= *a.n_dev)) return;
    const V3 o = v3(a.ox[ray], a.oy[ray], a.oz[ray]);
    const V3 d = v3(a.dx[ray], a.dy[ray], a.dz[ray]);
    const bool occl = a.all_occlusion != 0 || (a.flags != nullptr && (a.flags[ray] & 1u) != 0);
    TraceCount cnt;
    const TraceHit h = trace_one<COUNT, LDS_DEPTH, SPILL_DEPTH, NESTED>(a, o, d, a.time[ray], a.max_t[ray], occl,
                                                                 lds_stack + threadIdx.x, BLOCK, cnt);
    if (a.rec != nullptr) {
        const bool s = h.kind == PSY_HIT_SURFACE;
        uint4* r = a.rec + 2ull * ray;
        r[0] = make_uint4(s ? __float_as_uint(h.t) : 0u, s ? __float_as_uint(h.b0) : 0u,
                          s ? __float_as_uint(h.b1) : 0u, s ? __float_as_uint(h.b2) : 0u);
        r[1] = make_uint4(s ? h.prim : PSY_NONE, s ? h.inst : PSY_NONE, (uint32_t)h.kind, 0u);
    } else {
        a.kind[ray] = h.kind;
        if (h.kind == PSY_HIT_SURFACE) {
            a.t[ray] = h.t;
            a.b0[ray] = h.b0;
            a.b1[ray] = h.b1;
            a.b2[ray] = h.b2;
            a.prim_id[ray] = h.prim;
            a.inst_id[ray] = h.inst;
        }
    }
    if (COUNT) {
        atomicAdd(a.counters + 0, (unsigned long long)cnt.nodes);
        atomicAdd(a.counters + 1, (unsigned long long)cnt.mnodes);
        atomicAdd(a.counters + 2, (unsigned long long)cnt.tris);
        atomicAdd(a.counters + 3, (unsigned long long)cnt.mtris);
        atomicAdd(a.counters + 4, (unsigned long long)cnt.inst);
    }
}

}  // namespace

hipError_t launch_trace(const TraceArgs& args, bool counted, hipStream_t stream) {
    if (args.n == 0) return hipSuccess;
    uint64_t blocks = (args.n + BLOCK - 1) / BLOCK;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    // Scenes without sub-assemblies run the build with the nested walk
    // compiled out.
    const bool nested = args.groups != nullptr;
    auto k = counted ? (nested ? trace_kernel<true, true> : trace_kernel<true, false>)
                     : (nested ? trace_kernel<false, true> : trace_kernel<false, false>);
    hipLaunchKernelGGL(k, dim3((uint32_t)blocks), dim3(BLOCK), 0, stream, args);
    return hipGetLastError();
}

}  // namespace psy
