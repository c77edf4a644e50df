// Device-resident scene layout shared by the host packer and the gfx950
// kernels.  Everything here is plain data: it is written by the host builder
// (builder.cpp / scene.cpp) and read by the traversal kernels (trace.hip).
#pragma once

#include <stdint.h>

namespace psy {

// Child references inside a BVH4 node.
//   internal child: node index (bit 31 clear, bits 0..27)
//   leaf child:     bit 31 set, bits 2..27 first object, bits 0..1 count
//   absent child:   REF_EMPTY
// Mesh leaves hold 1..3 triangles (src/surface/triangle_mesh.rs:17); instance
// leaves hold exactly one instance (src/scene/assembly.rs:293).
// In the device node array bits 28..30 of the first two references (a node
// always has at least two children) carry the node's 6-bit traversal code,
// low three bits in child 0, so a static node's visit order comes from the
// order table without loading the node's last 16 bytes; device code masks
// them off (child_ref) before following a reference.
constexpr uint32_t REF_LEAF = 0x80000000u;
constexpr uint32_t REF_EMPTY = 0xffffffffu;
constexpr uint32_t REF_CODE_MASK = 0x70000000u;
constexpr uint32_t REF_CODE_SHIFT = 28;
constexpr uint32_t MAX_LEAF_FIRST = (1u << 26) - 1;
constexpr uint32_t MAX_NODES = 1u << 28;

static inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) { return REF_LEAF | (first << 2) | count; }

// One BVH4 node: a 64-B-aligned 128-B record.  The first 96 B are the four
// child boxes as SoA (BBox4, src/bbox4.rs:16-21) for time sample 0; nodes with
// time-sampled bounds keep all samples in the motion-bounds array instead.
//   order[o]  visit order for ray octant o, 2 bits per slot, first child in
//             the low bits (TRAVERSAL_TABLE[o][traversal_code],
//             sub_crates/bvh_order, src/accel/bvh4.rs:152-163).
//   meta      bits 0..2 child count, bits 8..15 bounds_len (time samples),
//             bits 16..21 traversal code.
//   motion    first BBox4 record of this node's time samples.
struct alignas(64) GpuNode {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
    uint8_t order[8];
    uint32_t meta;
    uint32_t motion;
};
static_assert(sizeof(GpuNode) == 128, "GpuNode must be 128 bytes");

// One time sample of a node's four child boxes (BBox4).
struct GpuBox4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
};
static_assert(sizeof(GpuBox4) == 96, "GpuBox4 must be 96 bytes");

// Triangle records: the three vertices in stored (winding-swapped) order
// (src/surface/triangle_mesh.rs:80) and the original triangle index.
// Static meshes use 48-B records; meshes with S deformation time samples use
// records of S*36+4 bytes rounded up to 16: vertex v of sample s at float
// offset s*9 + v*3, the index after all samples.
struct GpuTriStatic {
    float v[9];
    uint32_t prim;
    uint32_t pad[2];
};
static_assert(sizeof(GpuTriStatic) == 48, "GpuTriStatic must be 48 bytes");

static inline uint32_t tri_record_bytes(uint32_t time_samples) {
    return ((time_samples * 36u + 4u) + 15u) & ~15u;
}

// Per mesh: root reference of its BVH4, byte offset of its triangle records,
// record stride and time-sample count.
struct GpuMesh {
    uint32_t root;
    uint32_t time_samples;
    uint32_t tri_stride;   // bytes
    uint32_t rec_map;      // start of this mesh's triangle -> record map
    uint64_t tri_offset;   // bytes into the triangle buffer
    // Shading normals (triangle_mesh.rs:29-116, 272-296), PSY_NONE if none:
    // normal_base indexes the xyz normal array, vertex-major with the time
    // samples of a vertex together; nidx_base the per-triangle vertex
    // triples (stored order v0, v2, v1) indexed by original triangle.
    uint32_t normal_base;
    uint32_t nidx_base;
};

// Per instance (src/scene/assembly.rs:404-411).  `xf_start/xf_count` index
// the transform array (world -> object); `inv_static` indexes the inverse of
// the single transform of a non-motion instance (precomputed on the host with
// the same arithmetic the device uses).
struct GpuInstance {
    uint32_t mesh;
    uint32_t id;
    uint32_t shader;
    uint32_t xf_start;
    uint32_t xf_count;
    uint32_t inv_static;
    uint32_t kind;     // PSY_OBJECT_*; `mesh` indexes the light arrays for lights
    uint32_t pad;
};

// An instance path into an assembly of a nested scene (the assembly
// reached through a chain of assembly instances): the root of the
// assembly's own instance BVH, the first of this path's instance records
// (one per instance of the assembly, in its BVH's leaf order; sub-assembly
// instances are records of kind PSY_OBJECT_ASSEMBLY whose `mesh` names the
// path they open), and the transforms TransformStack::push composes into
// the assembly along the path (transform_stack.rs:33-60; none at the root
// assembly).  Traversal enters a sub-assembly by pushing a level marker on
// its node stack: a leaf reference with count 0 holding the enclosing path,
// restored when the marker is popped (tracer.rs:95-148).
struct GpuGroup {
    uint32_t root;
    uint32_t base;
    uint32_t xf_start;
    uint32_t xf_count;
};
constexpr uint32_t make_level_marker(uint32_t group) { return REF_LEAF | (group << 2); }

// Transform: glam Affine3A columns c0, c1, c2 and translation t, stored as
// three rows (x: c0.x c1.x c2.x t.x, ...) so each row is one float4.
struct GpuXform {
    float r[3][4];
};

// ---- renderer data ---------------------------------------------------------

// Color (src/color.rs:46-62): kind PSY_COLOR_*, then X,Y,Z or
// (temperature, factor, unused).
struct GpuColor {
    uint32_t kind;
    float a, b, c;
};

// Lights.  A light with time samples keeps them in a float array beside the
// records: `samples` is the offset of its lists there and `counts` their
// lengths (one byte each, 0 = the value in the record); each list is lerped
// on its own at the ray's time.  Rectangles: (w, h) pairs, then colours;
// spheres: radii, then colours; distant lights: radii, directions (xyz),
// then colours.  A colour sample is four floats: kind (0, 1 or 2), a, b, c.
struct GpuRectLight {
    float w, h;
    uint32_t samples, counts;  // counts: dims | colours << 8
    GpuColor color;
};
struct GpuSphereLight {
    float radius;
    uint32_t samples, counts;  // counts: radii | colours << 8
    float pad;
    GpuColor color;
};
struct GpuDistantLight {
    float radius;
    float dir[3];
    GpuColor color;
    uint32_t samples, counts;  // counts: radii | directions << 8 | colours << 16
    float energy;              // approximate_energy, the mean over the colours
    float pad;
};

// A surface closure as shaded (SimpleSurfaceShader::shade, src/shading/mod.rs:
// 42-61): kind PSY_SHADER_*.
struct GpuClosure {
    uint32_t kind;
    float roughness, fresnel;
    float ior;  // PSY_SHADER_DIELECTRIC only
    GpuColor color;
};

// Light tree node (src/accel/light_tree.rs:17-33) after the collapse to
// arity 8: children are contiguous nodes [child_first, child_first +
// child_count); a leaf (child_count 0) names an entry of the light list.
// Bounds are time samples [bounds_start, bounds_start + bounds_len) of the
// light-bounds array.
struct GpuLightNode {
    uint32_t child_first;
    uint32_t child_count;
    uint32_t light;
    uint32_t bounds_start;
    uint32_t bounds_len;
    float energy;
    uint32_t pad[2];
};

// A box of the light tree: min xyz, max xyz.
struct GpuBox {
    float lo[3], hi[3];
};

// A light-tree leaf (GpuLightNode::light of a node without children).  Each
// assembly on an instance path has its own tree (assembly.rs:297-338); a
// leaf is a light instance or a sub-assembly instance whose tree continues
// the selection (assembly.rs:46-141).  Transform lists live in the light
// transform array:
//   kind 0 (light): target = the instance's leaf position; top = the
//     enclosing assemblies' composed transforms (the transform stack's top),
//     own = the light instance's own transforms; the light is sampled with
//     lerp(top) * lerp(own).
//   kind 1 (sub-assembly): target = the root node of its tree; top = the
//     composed transforms into it, whose lerp takes the shading point into
//     the sub-assembly's space.
struct GpuLightLeaf {
    uint32_t kind;
    uint32_t target;
    uint32_t top_start, top_count;
    uint32_t own_start, own_count;
    uint32_t pad[2];
};
constexpr uint32_t LIGHT_LEAF_LIGHT = 0u;
constexpr uint32_t LIGHT_LEAF_ASSEMBLY = 1u;
// Light trees chain through at most this many assemblies (the root's and
// three levels of sub-assemblies holding lights).
constexpr uint32_t MAX_LIGHT_LEVELS = 4u;

}  // namespace psy
