// Host-side scene packing: meshes and instances (the reference's
// TriangleMesh and AssemblyBuilder) flattened into the device layout.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/psy_rt.h"
#include "builder.h"
#include "layout.h"

namespace psy {

// The whole scene in device layout, still in host memory.
struct PackedScene {
    std::vector<GpuNode> nodes;
    std::vector<GpuBox4> motion;
    std::vector<uint8_t> tris;         // triangle records, per-mesh stride
    std::vector<GpuMesh> meshes;
    std::vector<GpuInstance> instances;  // in instance-BVH leaf order (per path of a nested scene)
    std::vector<GpuGroup> groups;        // nested scenes: instance paths into assemblies, the root's first
    std::vector<GpuXform> xforms;
    std::vector<GpuXform> inv_static;
    std::vector<float> normals;        // xyz per (vertex, time sample), see GpuMesh
    std::vector<uint32_t> nidx;        // stored-order vertex triples per triangle
    std::vector<uint8_t> order_table;  // TRAVERSAL_TABLE[octant][code], 8 x 48 bytes
    uint32_t top_root = REF_EMPTY;

    // Lights and shading lookups.
    std::vector<GpuRectLight> rect_lights;
    std::vector<GpuSphereLight> sphere_lights;
    std::vector<float> light_samples;       // time samples of the lights, see GpuRectLight
    std::vector<uint32_t> inst_pos;         // description index -> leaf position
    std::vector<uint32_t> tri_rec;          // per mesh at GpuMesh::rec_map: triangle -> record
    std::vector<GpuLightNode> light_nodes;  // light trees, the root assembly's root first; empty without lights
    std::vector<GpuBox> light_boxes;
    std::vector<GpuLightLeaf> light_leaves; // leaves of the light trees
    std::vector<GpuXform> light_xforms;     // transform lists of the leaves

    // Host builds, kept for parity dumps (index = mesh; the instance BVH last).
    std::vector<HostBvh4> bvhs;
    bool has_xf = false;        // some instance has transforms
    bool has_motion = false;    // time-sampled node bounds or vertices
    bool has_lights = false;    // some instance places a light
    bool has_normals = false;   // some mesh has shading normals
    bool nested = false;        // sub-assemblies are traversed as nested BVHs (`groups`)
    uint32_t max_stack = 0;     // deepest pending-node count traversal can need
    uint32_t max_depth = 0;
    uint64_t leaf_count = 0;
    uint64_t tri_count = 0;
    double build_seconds = 0;
    uint32_t device_built = 0;      // mesh BVHs built by the device builder
    uint32_t device_light_trees = 0;  // light trees partitioned on the device
    double device_build_ms = 0;     // their build time (uploads and downloads included)
};

// Where the mesh BVHs are built: on the host, or on the current HIP device
// for meshes of at least `device_min_tris` triangles (the device builder
// builds the host builder's trees bit for bit; a mesh it cannot take, or a
// HIP error, falls back to the host builder).
struct BuildOptions {
    bool device = false;
    uint32_t device_min_tris = 0;
};

// Validates `desc` and builds every BVH.  Returns PSY_OK or an error code
// with a message in `err`.
int pack_scene(const psy_scene_desc* desc, PackedScene& out, std::string& err, const BuildOptions& opt = {});

// TransformStack::push (src/transform_stack.rs:33-60): the parent's and an
// instance's time samples combined by merge_slices_to (src/algorithm.rs:
// 266-305), the shorter list resampled with lerp_slice, each pair composed
// as math3d's parent * child.
std::vector<Xf> push_xforms(const std::vector<Xf>& top, const std::vector<Xf>& own);

// Light time samples (shared with the renderer's distant lights): the mean
// colour energy, a checked colour list, and the packing of several samples
// into a light sample array (see GpuRectLight).
float light_energy(const std::vector<psy_color>& colours);
bool colour_list(const psy_color& one, uint32_t n, const psy_color* cs, std::vector<psy_color>& out,
                 std::string& why);
uint32_t pack_samples(uint32_t n, const float* list, uint32_t width, float* one, std::vector<float>& samples);
uint32_t pack_colours(const std::vector<psy_color>& cs, std::vector<float>& samples);

}  // namespace psy
