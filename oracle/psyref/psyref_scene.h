// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Scene model and tracer: restates src/surface/triangle_mesh.rs,
// src/surface/mod.rs (hit record), src/scene/assembly.rs (instances and the
// instance BVH), src/transform_stack.rs and src/tracer.rs.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "psyref_bvh.h"
#include "psyref_lights.h"
#include "psyref_tri.h"

namespace psyref {

// surface/mod.rs:70-92.  The reference's record carries no primitive or
// instance id; prim_id (the triangle's original index, triangle_mesh.rs:80)
// and inst_id (Instance::id, scene/assembly.rs:256) are added so hits can be
// compared one to one.
enum IsectKind : uint8_t { ISECT_MISS = 0, ISECT_OCCLUDE = 1, ISECT_HIT = 2 };
struct Isect {
    uint8_t kind = ISECT_MISS;
    float t = 0, b0 = 0, b1 = 0, b2 = 0;
    uint32_t prim_id = 0xffffffffu, inst_id = 0xffffffffu;
    int32_t shader = -1;  // surface shader index of the instance (-1: unassigned)
    Vec3 incoming, pos, nor, nor_g;
    float pos_err = 0;
    Transform local_space;
    float sample_pdf = 0;
    bool light = false;      // hit a light surface: Emit(emit) closure
    Color emit;
};

// triangle_mesh.rs:19-116
struct TriangleMesh {
    static const size_t MAX_LEAF_TRIANGLE_COUNT = 3;  // triangle_mesh.rs:17
    size_t time_sample_count = 1;
    std::vector<Vec3> vertices;   // [vert][time]
    std::vector<Vec3> normals;    // empty or same layout
    struct Tri { uint32_t v0, v1, v2, orig; };
    std::vector<Tri> indices;     // winding-swapped, BVH order
    BVH4 accel;

    // verts: [time][vert][3]; normals optional, same layout; tris: [n][3]
    static std::unique_ptr<TriangleMesh> from_verts_and_indices(const float* verts, size_t vert_count,
                                                                size_t time_samples, const float* normals,
                                                                const uint32_t* tris, size_t tri_count);
    const std::vector<BBox>& bounds() const { return accel.root_bounds; }

    void intersect_rays(RayBatch& rays, RayStack& ray_stack, std::vector<Isect>& isects, int32_t shader,
                        uint32_t inst_id, const Transform* space, size_t space_len, uint64_t& node_tests) const;
};

// transform_stack.rs:8-83
struct TransformStack {
    std::vector<Transform> stack;
    std::vector<size_t> stack_indices{0, 0};
    void clear() { stack.clear(); stack_indices.assign({0, 0}); }
    void push(const Transform* xf, size_t n);
    void pop();
    const Transform* top(size_t& n) const {
        size_t sil = stack_indices.size();
        size_t i1 = stack_indices[sil - 2], i2 = stack_indices[sil - 1];
        n = i2 - i1;
        return stack.data() + i1;
    }
};

// scene/assembly.rs:398-417.  Objects are meshes (Object::Surface) or
// light surfaces (Object::SurfaceLight).
enum ObjectKind : uint8_t { OBJ_MESH = 0, OBJ_RECT_LIGHT = 1, OBJ_SPHERE_LIGHT = 2 };
struct Instance {
    uint8_t obj_kind = OBJ_MESH;
    bool is_assembly = false;
    size_t data_index = 0;
    int32_t shader = -1;
    size_t id = 0;
    bool has_xforms = false;
    size_t xf_start = 0, xf_end = 0;
};

struct Assembly {
    std::vector<Instance> instances;
    std::vector<Transform> xforms;
    std::vector<std::unique_ptr<TriangleMesh>> objects;  // surface objects
    std::vector<RectangleLight> rect_lights;
    std::vector<SphereLight> sphere_lights;
    std::vector<std::unique_ptr<Assembly>> assemblies;
    BVH4 object_accel;
    // Per-instance world bounds, indexed by Instance::id (assembly.rs:354-395)
    std::vector<size_t> inst_bound_start;
    std::vector<BBox> inst_bounds;
    const std::vector<BBox>& bounds() const { return object_accel.root_bounds; }

    // AssemblyBuilder::add_instance (assembly.rs:223-282)
    void add_instance(bool is_assembly, size_t data_index, int32_t shader, const Transform* xf, size_t n_xf,
                      uint8_t obj_kind = OBJ_MESH);
    std::vector<BBox> object_bounds(const Instance& inst) const;
    float light_energy(const Instance& inst) const;
    // AssemblyBuilder::build (assembly.rs:288-395), object accel only.
    void build();
};

// tracer.rs:15-191
struct Tracer {
    const Assembly* root;
    uint64_t ray_trace_count = 0;
    uint64_t node_tests = 0;
    RayStack ray_stack;
    TransformStack xform_stack;
    std::vector<Isect> isects;

    explicit Tracer(const Assembly* a) : root(a) {}
    const std::vector<Isect>& trace(RayBatch& rays);

   private:
    void trace_assembly(const Assembly* assembly, RayBatch& rays);
};

}  // namespace psyref
