// Launch interface of the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/psy_rt.h"
#include "layout.h"

namespace psy {

constexpr int TRACE_BLOCK = 128;           // lanes per workgroup (two wave64s)
constexpr int TRACE_LDS_DEPTH = 24;        // stack entries per lane held in LDS
// Total stack entries per lane: the bound 3 * (instance-BVH depth + mesh-BVH
// depth) + 2 for two BVH4s of the deepest the builder makes (depth 22, from
// a binary depth of at most 42: src/accel/bvh_base.rs:7, bvh4.rs:87).
constexpr int TRACE_STACK_CAPACITY = 3 * (22 + 22) + 2;

struct TraceArgs {
    // scene
    const GpuNode* nodes;
    const GpuBox4* motion;
    const uint8_t* tris;
    const GpuMesh* meshes;
    const GpuInstance* instances;
    const GpuXform* xforms;
    const GpuXform* inv_static;
    const GpuRectLight* rect_lights;
    const GpuSphereLight* sphere_lights;
    const float* light_samples;  // time samples of the lights (layout.h)
    const uint8_t* order_table;  // TRAVERSAL_TABLE[octant][code], 8 x 48 bytes
    const GpuGroup* groups;      // nested scenes: instance paths (layout.h), else NULL
    uint32_t top_root;
    // rays (device)
    const float *ox, *oy, *oz, *dx, *dy, *dz, *time, *max_t;
    const uint8_t* flags;
    uint32_t all_occlusion;    // every ray is an occlusion ray (flags ignored)
    uint64_t n;                // ray count, or the capacity when n_dev is set
    const uint32_t* n_dev;     // device-resident ray count (persistent kernel), or NULL
    // hits (device): 32-byte records when `rec` is set, else structure of
    // arrays
    uint4* rec;
    uint8_t* kind;
    float *t, *b0, *b1, *b2;
    uint32_t *prim_id, *inst_id;
    // counters (device, 5 x u64) for the instrumented kernel
    unsigned long long* counters;
};

// Launch shape of the persistent kernel (chosen per scene by the autotuner).
struct TraceTuning {
    uint32_t waves;       // wave64 workgroups launched
    uint32_t refill_min;  // idle lanes that trigger a refill (1..64)
    uint32_t lds_depth;   // stack entries per lane kept in LDS (8, 16, 24)
    uint32_t occupancy;   // register budget as waves per SIMD (4 or 8)
    uint32_t leaf_min;    // waiting lanes that end a descent round (1..64), 0: per-lane descents
    uint32_t coop;        // 1: wave-cooperative node fetch (static scenes)
};

// One lane per ray, one launch-wide grid (trace.hip).
hipError_t launch_trace(const TraceArgs& args, bool counted, hipStream_t stream);

// Persistent waves with lane refill (trace_persistent.hip).  `work` is a
// device counter the launch resets.
hipError_t launch_trace_persistent(const TraceArgs& args, bool counted, bool has_xf, bool has_motion,
                                   unsigned int* work, const TraceTuning& tune, hipStream_t stream);

}  // namespace psy
