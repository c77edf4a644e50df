// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
#include "psyref_scene.h"

#include <algorithm>
#include <cassert>

namespace psyref {

// ---- triangle_mesh.rs:29-116 -----------------------------------------------
std::unique_ptr<TriangleMesh> TriangleMesh::from_verts_and_indices(const float* verts, size_t vert_count,
                                                                   size_t time_samples, const float* nrm,
                                                                   const uint32_t* tris, size_t tri_count) {
    auto m = std::make_unique<TriangleMesh>();
    m->time_sample_count = time_samples;
    auto vtx = [&](const float* base, size_t ti, size_t vi) {
        const float* p = base + (ti * vert_count + vi) * 3;
        return Vec3(p[0], p[1], p[2]);
    };
    // Time samples of each vertex stored contiguously (:40-52, :56-72).
    m->vertices.resize(vert_count * time_samples);
    for (size_t vi = 0; vi < vert_count; vi++)
        for (size_t ti = 0; ti < time_samples; ti++) m->vertices[vi * time_samples + ti] = vtx(verts, ti, vi);
    if (nrm) {
        m->normals.resize(vert_count * time_samples);
        for (size_t vi = 0; vi < vert_count; vi++)
            for (size_t ti = 0; ti < time_samples; ti++) m->normals[vi * time_samples + ti] = vtx(nrm, ti, vi);
    }
    // Winding swap: stored as (v0, v2, v1, original index)  (:74-83)
    m->indices.resize(tri_count);
    for (size_t i = 0; i < tri_count; i++)
        m->indices[i] = Tri{tris[i * 3 + 0], tris[i * 3 + 2], tris[i * 3 + 1], (uint32_t)i};
    // Per-triangle time-sampled bounds, from the original vertex order (:85-100)
    std::vector<BBox> bounds;
    bounds.reserve(tri_count * time_samples);
    for (size_t i = 0; i < tri_count; i++)
        for (size_t ti = 0; ti < time_samples; ti++) {
            Vec3 p0 = vtx(verts, ti, tris[i * 3 + 0]);
            Vec3 p1 = vtx(verts, ti, tris[i * 3 + 1]);
            Vec3 p2 = vtx(verts, ti, tris[i * 3 + 2]);
            bounds.push_back(BBox(vmin(p0, vmin(p1, p2)), vmax(p0, vmax(p1, p2))));
        }
    // BVH over triangle keys (the original index) (:102-106)
    std::vector<uint32_t> keys(tri_count);
    for (size_t i = 0; i < tri_count; i++) keys[i] = (uint32_t)i;
    const BBox* bptr = bounds.data();
    m->accel = BVH4::from_objects(keys.data(), tri_count, MAX_LEAF_TRIANGLE_COUNT, [&](uint32_t k) {
        return BoundsSlice{bptr + (size_t)k * time_samples, time_samples};
    });
    // The reference permutes the index tuples themselves; apply the same
    // permutation to the tuple array.
    std::vector<Tri> permuted(tri_count);
    for (size_t i = 0; i < tri_count; i++) permuted[i] = m->indices[keys[i]];
    m->indices.swap(permuted);
    return m;
}

// ---- triangle_mesh.rs:118-322 ----------------------------------------------
void TriangleMesh::intersect_rays(RayBatch& rays, RayStack& ray_stack, std::vector<Isect>& isects, int32_t shader,
                                  uint32_t inst_id, const Transform* space, size_t space_len,
                                  uint64_t& node_tests) const {
    // Static transform for the non-motion case (:128-139)
    Transform static_mat_space = space_len == 1 ? xform_inverse(lerp_slice(space, space_len, 0.0f)) : Transform();

    accel.traverse(
        rays, ray_stack,
        [&](size_t obj_start, size_t obj_end) {
            size_t tri_count = obj_end - obj_start;
            bool is_cached =
                ray_stack.ray_count_in_next_task() >= tri_count && time_sample_count == 1 && space_len <= 1;
            Vec3 cache[MAX_LEAF_TRIANGLE_COUNT][3];
            if (is_cached) {
                for (size_t ti = obj_start; ti < obj_end; ti++) {
                    const Tri& tr = indices[ti];
                    Vec3* c = cache[ti - obj_start];
                    c[0] = vertices[tr.v0];
                    c[1] = vertices[tr.v1];
                    c[2] = vertices[tr.v2];
                    if (space_len != 0) {
                        c[0] = xform_point(c[0], static_mat_space);
                        c[1] = xform_point(c[1], static_mat_space);
                        c[2] = xform_point(c[2], static_mat_space);
                    }
                }
            }

            ray_stack.do_next_task([&](size_t ray_idx) {
                if (rays.is_done(ray_idx)) return;
                float ray_time = rays.time(ray_idx);
                Transform mat_space =
                    space_len > 1 ? xform_inverse(lerp_slice(space, space_len, ray_time)) : static_mat_space;

                bool non_shadow_hit = false;
                Vec3 hit_tri[3];
                Tri hit_idx{};
                TriHit hit_data{};
                RayTriPrecompute pre(rays.dir(ray_idx));
                for (size_t ti = obj_start; ti < obj_end; ti++) {
                    const Tri& tr = indices[ti];
                    Vec3 tri[3];
                    if (is_cached) {
                        tri[0] = cache[ti - obj_start][0];
                        tri[1] = cache[ti - obj_start][1];
                        tri[2] = cache[ti - obj_start][2];
                    } else {
                        if (time_sample_count == 1) {
                            tri[0] = vertices[tr.v0];
                            tri[1] = vertices[tr.v1];
                            tri[2] = vertices[tr.v2];
                        } else {
                            size_t S = time_sample_count;
                            tri[0] = lerp_slice(&vertices[(size_t)tr.v0 * S], S, ray_time);
                            tri[1] = lerp_slice(&vertices[(size_t)tr.v1 * S], S, ray_time);
                            tri[2] = lerp_slice(&vertices[(size_t)tr.v2 * S], S, ray_time);
                        }
                        if (space_len != 0) {
                            tri[0] = xform_point(tri[0], mat_space);
                            tri[1] = xform_point(tri[1], mat_space);
                            tri[2] = xform_point(tri[2], mat_space);
                        }
                    }
                    TriHit h;
                    if (intersect_triangle(rays.orig(ray_idx), pre, rays.max_t(ray_idx), tri[0], tri[1], tri[2], h)) {
                        if (rays.is_occlusion(ray_idx)) {
                            Isect& is = isects[ray_idx];
                            is = Isect();
                            is.kind = ISECT_OCCLUDE;
                            is.t = h.t; is.b0 = h.b0; is.b1 = h.b1; is.b2 = h.b2;
                            is.prim_id = tr.orig;
                            is.inst_id = inst_id;
                            rays.mark_done(ray_idx);
                            break;
                        } else {
                            non_shadow_hit = true;
                            rays.set_max_t(ray_idx, h.t);
                            hit_tri[0] = tri[0]; hit_tri[1] = tri[1]; hit_tri[2] = tri[2];
                            hit_idx = tr;
                            hit_data = h;
                        }
                    }
                }

                if (non_shadow_hit) {
                    Isect& is = isects[ray_idx];
                    is = Isect();
                    is.kind = ISECT_HIT;
                    is.t = hit_data.t; is.b0 = hit_data.b0; is.b1 = hit_data.b1; is.b2 = hit_data.b2;
                    is.prim_id = hit_idx.orig;
                    is.inst_id = inst_id;
                    is.shader = shader;
                    surface_point(hit_tri[0], hit_tri[1], hit_tri[2], hit_data.b0, hit_data.b1, hit_data.b2, is.pos,
                                  is.pos_err);
                    Vec3 geo_normal = cross(hit_tri[0] - hit_tri[1], hit_tri[0] - hit_tri[2]);
                    Vec3 shading_normal = geo_normal;
                    if (!normals.empty()) {
                        size_t S = time_sample_count;
                        Vec3 n0 = normalized(lerp_slice(&normals[(size_t)hit_idx.v0 * S], S, ray_time));
                        Vec3 n1 = normalized(lerp_slice(&normals[(size_t)hit_idx.v1 * S], S, ray_time));
                        Vec3 n2 = normalized(lerp_slice(&normals[(size_t)hit_idx.v2 * S], S, ray_time));
                        Vec3 s_nor = xform_normal(((n0 * hit_data.b0) + (n1 * hit_data.b1)) + (n2 * hit_data.b2),
                                                  mat_space);
                        shading_normal = dot(s_nor, geo_normal) >= 0.0f ? s_nor : neg(s_nor);
                    }
                    is.incoming = rays.dir(ray_idx);
                    is.nor = shading_normal;
                    is.nor_g = geo_normal;
                    is.local_space = mat_space;
                    is.sample_pdf = 0.0f;
                }
            });
            ray_stack.pop_task();
        },
        node_tests);
}

// ---- transform_stack.rs ----------------------------------------------------
void TransformStack::push(const Transform* xf, size_t n) {
    assert(n > 0);
    if (stack.empty()) {
        stack.insert(stack.end(), xf, xf + n);
    } else {
        size_t sil = stack_indices.size();
        size_t i1 = stack_indices[sil - 2], i2 = stack_indices[sil - 1];
        size_t n1 = i2 - i1;
        std::vector<Transform> parent(stack.begin() + i1, stack.begin() + i2);
        // merge_slices_to (algorithm.rs:266-305) with merge = parent * child
        if (n1 == n) {
            for (size_t i = 0; i < n; i++) stack.push_back(xform_compose(parent[i], xf[i]));
        } else if (n1 > n) {
            float s = (float)(n1 - 1);
            for (size_t i = 0; i < n1; i++) stack.push_back(xform_compose(parent[i], lerp_slice(xf, n, (float)i / s)));
        } else {
            float s = (float)(n - 1);
            for (size_t i = 0; i < n; i++)
                stack.push_back(xform_compose(lerp_slice(parent.data(), n1, (float)i / s), xf[i]));
        }
    }
    stack_indices.push_back(stack.size());
}

void TransformStack::pop() {
    assert(stack_indices.size() > 2);
    size_t sil = stack_indices.size();
    size_t i1 = stack_indices[sil - 2], i2 = stack_indices[sil - 1];
    stack.resize(stack.size() - (i2 - i1));
    stack_indices.pop_back();
}

// ---- scene/assembly.rs -----------------------------------------------------
void Assembly::add_instance(bool is_assembly, size_t data_index, int32_t shader, const Transform* xf, size_t n_xf,
                            uint8_t obj_kind) {
    Instance inst;
    inst.obj_kind = obj_kind;
    inst.is_assembly = is_assembly;
    inst.data_index = data_index;
    inst.shader = shader;
    inst.id = instances.size();
    if (xf && n_xf > 0) {
        inst.has_xforms = true;
        inst.xf_start = xforms.size();
        inst.xf_end = xforms.size() + n_xf;
        xforms.insert(xforms.end(), xf, xf + n_xf);
    }
    instances.push_back(inst);
}

void Assembly::build() {
    // instance_bounds (assembly.rs:354-395)
    std::vector<size_t> bis{0};
    std::vector<BBox> bbs;
    for (const Instance& inst : instances) {
        std::vector<BBox> src = object_bounds(inst);
        std::vector<BBox> out;
        if (inst.has_xforms)
            transform_bbox_slice_from(src.data(), src.size(), xforms.data() + inst.xf_start,
                                      inst.xf_end - inst.xf_start, out);
        else
            out = src;
        bbs.insert(bbs.end(), out.begin(), out.end());
        bis.push_back(bbs.size());
    }
    // Object accel with one instance per leaf (assembly.rs:292-295); the
    // instances array is permuted by the build.
    std::vector<uint32_t> keys(instances.size());
    for (size_t i = 0; i < instances.size(); i++) keys[i] = (uint32_t)i;
    const BBox* bp = bbs.data();
    object_accel = BVH4::from_objects(keys.data(), keys.size(), 1, [&](uint32_t k) {
        size_t id = instances[k].id;
        return BoundsSlice{bp + bis[id], bis[id + 1] - bis[id]};
    });
    std::vector<Instance> permuted(instances.size());
    for (size_t i = 0; i < keys.size(); i++) permuted[i] = instances[keys[i]];
    instances.swap(permuted);
    inst_bound_start = bis;
    inst_bounds = bbs;
}

std::vector<BBox> Assembly::object_bounds(const Instance& inst) const {
    if (inst.is_assembly) return assemblies[inst.data_index]->bounds();
    switch (inst.obj_kind) {
        case OBJ_RECT_LIGHT: return rect_lights[inst.data_index].bounds_list();
        case OBJ_SPHERE_LIGHT: return sphere_lights[inst.data_index].bounds_list();
        default: return objects[inst.data_index]->bounds();
    }
}

float Assembly::light_energy(const Instance& inst) const {
    switch (inst.obj_kind) {
        case OBJ_RECT_LIGHT: return rect_lights[inst.data_index].approximate_energy();
        case OBJ_SPHERE_LIGHT: return sphere_lights[inst.data_index].approximate_energy();
        default: return 0.0f;
    }
}

// Light surfaces as trace leaves: RectangleLight::intersect_rays
// (rectangle_light.rs:257-338) and SphereLight::intersect_rays
// (sphere_light.rs:206-339).
static void intersect_light_rays(const Assembly* a, const Instance& inst, RayBatch& rays, RayStack& ray_stack,
                                 std::vector<Isect>& isects, const Transform* space, size_t space_len) {
    ray_stack.pop_do_next_task([&](size_t ri) {
        float time = rays.time(ri);
        Transform xform = lerp_slice(space, space_len, time);
        float t, pos_err, pdf;
        Vec3 pos, nrm;
        bool hit;
        Color emit;
        if (inst.obj_kind == OBJ_RECT_LIGHT) {
            const RectangleLight l = a->rect_lights[inst.data_index].at(time);
            hit = l.intersect(xform, rays.orig(ri), rays.dir(ri), rays.max_t(ri), t, pos, pos_err, nrm, pdf);
            emit = l.emit_color();
        } else {
            const SphereLight l = a->sphere_lights[inst.data_index].at(time);
            hit = l.intersect(xform, rays.orig(ri), rays.orig_local(ri), rays.dir(ri), rays.max_t(ri), t, pos,
                              pos_err, nrm, pdf);
            emit = l.emit_color();
        }
        if (!hit) return;
        Isect& is = isects[ri];
        if (rays.is_occlusion(ri)) {
            is = Isect();
            is.kind = ISECT_OCCLUDE;
            is.inst_id = (uint32_t)inst.id;
            rays.mark_done(ri);
            return;
        }
        is = Isect();
        is.kind = ISECT_HIT;
        is.t = t;
        is.inst_id = (uint32_t)inst.id;
        is.incoming = rays.dir(ri);
        is.pos = pos;
        is.pos_err = pos_err;
        is.nor = nrm;
        is.nor_g = nrm;
        is.local_space = xform;
        is.sample_pdf = pdf;
        is.light = true;
        is.emit = emit;
        rays.set_max_t(ri, t);
    });
}

// ---- tracer.rs -------------------------------------------------------------
const std::vector<Isect>& Tracer::trace(RayBatch& rays) {
    ray_trace_count += rays.size();
    ray_stack.clear();
    isects.assign(rays.size(), Isect());
    Transform ident;
    for (size_t i = 0; i < rays.size(); i++) rays.update_local(i, ident);
    ray_stack.ensure_lane_count(8);
    for (size_t i = 0; i < rays.size(); i++) ray_stack.push_ray_index(i, ray_code(rays.dir(i)));
    const size_t all_lanes[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    ray_stack.push_lanes_to_tasks(all_lanes, 8);
    while (!ray_stack.is_empty()) trace_assembly(root, rays);
    return isects;
}

void Tracer::trace_assembly(const Assembly* assembly, RayBatch& rays) {
    assembly->object_accel.traverse(
        rays, ray_stack,
        [&](size_t obj_start, size_t /*obj_end*/) {
            const Instance& inst = assembly->instances[obj_start];
            if (inst.has_xforms) {
                xform_stack.push(assembly->xforms.data() + inst.xf_start, inst.xf_end - inst.xf_start);
                size_t n;
                const Transform* xf = xform_stack.top(n);
                ray_stack.do_next_task([&](size_t ri) { rays.update_local(ri, lerp_slice(xf, n, rays.time(ri))); });
                ray_stack.duplicate_next_task();
            }
            if (!inst.is_assembly && inst.obj_kind != OBJ_MESH) {
                size_t n;
                const Transform* xf = xform_stack.top(n);
                intersect_light_rays(assembly, inst, rays, ray_stack, isects, xf, n);
            } else if (!inst.is_assembly) {
                size_t n;
                const Transform* xf = xform_stack.top(n);
                assembly->objects[inst.data_index]->intersect_rays(rays, ray_stack, isects, inst.shader,
                                                                   (uint32_t)inst.id, xf, n, node_tests);
            } else {
                trace_assembly(assembly->assemblies[inst.data_index].get(), rays);
            }
            if (inst.has_xforms) {
                xform_stack.pop();
                size_t n;
                const Transform* xf = xform_stack.top(n);
                if (n != 0) {
                    ray_stack.pop_do_next_task(
                        [&](size_t ri) { rays.update_local(ri, lerp_slice(xf, n, rays.time(ri))); });
                } else {
                    Transform ident;
                    ray_stack.pop_do_next_task([&](size_t ri) { rays.update_local(ri, ident); });
                }
            }
        },
        node_tests);
}

}  // namespace psyref
