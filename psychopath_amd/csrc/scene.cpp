// Scene packing: TriangleMesh::from_verts_and_indices
// (src/surface/triangle_mesh.rs:29-116) and AssemblyBuilder::build
// (src/scene/assembly.rs:288-395) restated over the device layout.
#include "scene.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace psy {

namespace {

constexpr uint32_t MAX_TIME_SAMPLES = 64;
constexpr uint32_t MESH_LEAF_TRIS = 3;  // triangle_mesh.rs:17

Xf xf_at(const psy_scene_desc* d, uint32_t i) { return xf_from_values(d->xforms + 12ull * i); }

Xf lerp_slice_xf(const Xf* s, uint32_t n, float alpha) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(n, alpha, i1, a2)) return s[n - 1];
    return xf_lerp(s[i1], s[i1 + 1], a2);
}
Box lerp_slice_bx(const Box* s, uint32_t n, float alpha) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(n, alpha, i1, a2)) return s[n - 1];
    return box_lerp(s[i1], s[i1 + 1], a2);
}

// transform_bbox_slice_from (src/bbox.rs:148-169): object-space bounds to
// world space through the inverses of the instance's world-to-object
// transforms, resampling the shorter list.
std::vector<Box> transform_bounds(const std::vector<Box>& in, const Xf* xf, uint32_t nx) {
    std::vector<Box> out;
    uint32_t nb = (uint32_t)in.size();
    if (nx == 0) return in;
    if (nb == nx) {
        for (uint32_t i = 0; i < nb; i++) out.push_back(box_transformed(in[i], xf_inverse(xf[i])));
    } else if (nb > nx) {
        float s = (float)(nb - 1);
        for (uint32_t i = 0; i < nb; i++)
            out.push_back(box_transformed(in[i], xf_inverse(lerp_slice_xf(xf, nx, (float)i / s))));
    } else {
        float s = (float)(nx - 1);
        for (uint32_t i = 0; i < nx; i++)
            out.push_back(box_transformed(lerp_slice_bx(in.data(), nb, (float)i / s), xf_inverse(xf[i])));
    }
    return out;
}

GpuXform to_gpu(const Xf& x) {
    GpuXform g;
    std::memcpy(g.r, x.m, sizeof(g.r));
    return g;
}

// Appends `bvh` to the scene arrays, rebasing node and motion indices.
uint32_t append_bvh(PackedScene& s, const HostBvh4& bvh) {
    uint32_t node_base = (uint32_t)s.nodes.size();
    uint32_t motion_base = (uint32_t)s.motion.size();
    for (GpuNode n : bvh.nodes) {
        uint32_t cnt = n.meta & 7u;
        for (uint32_t c = 0; c < cnt; c++)
            if (!(n.child[c] & REF_LEAF)) n.child[c] += node_base;
        if (n.motion != REF_EMPTY) n.motion += motion_base;
        s.nodes.push_back(n);
    }
    s.motion.insert(s.motion.end(), bvh.motion.begin(), bvh.motion.end());
    if (bvh.root == REF_EMPTY || (bvh.root & REF_LEAF)) return bvh.root;
    return bvh.root + node_base;
}

}  // namespace

int pack_scene(const psy_scene_desc* d, PackedScene& s, std::string& err) {
    auto t0 = std::chrono::steady_clock::now();
    if (!d) { err = "scene description is NULL"; return PSY_E_INVALID; }
    if (d->mesh_count > 0 && !d->meshes) { err = "meshes is NULL"; return PSY_E_INVALID; }
    if (d->instance_count > 0 && !d->instances) { err = "instances is NULL"; return PSY_E_INVALID; }
    if (d->xform_count > 0 && !d->xforms) { err = "xforms is NULL"; return PSY_E_INVALID; }
    if (d->instance_count > MAX_LEAF_FIRST) { err = "too many instances"; return PSY_E_RANGE; }

    s.bvhs.resize(d->mesh_count + 1);
    s.meshes.resize(d->mesh_count);

    // ---- meshes -----------------------------------------------------------
    for (uint32_t m = 0; m < d->mesh_count; m++) {
        const psy_mesh_desc& md = d->meshes[m];
        if (!md.verts || !md.indices || md.vert_count == 0 || md.tri_count == 0) {
            err = "mesh " + std::to_string(m) + ": empty or NULL arrays";
            return PSY_E_INVALID;
        }
        if (md.time_samples == 0 || md.time_samples > MAX_TIME_SAMPLES) {
            err = "mesh " + std::to_string(m) + ": time_samples must be in [1, 64]";
            return PSY_E_RANGE;
        }
        if (md.tri_count > MAX_LEAF_FIRST) { err = "mesh " + std::to_string(m) + ": too many triangles"; return PSY_E_RANGE; }
        const uint32_t S = md.time_samples, V = md.vert_count, T = md.tri_count;
        for (uint64_t i = 0; i < 3ull * T; i++)
            if (md.indices[i] >= V) {
                err = "mesh " + std::to_string(m) + ": vertex index out of range";
                return PSY_E_INVALID;
            }
        auto vtx = [&](uint32_t ti, uint32_t vi) {
            const float* p = md.verts + (3ull * ((uint64_t)ti * V + vi));
            return V3{p[0], p[1], p[2]};
        };
        // Per-triangle time-sampled bounds in the file's vertex order
        // (triangle_mesh.rs:85-100).
        std::vector<Box> bounds((uint64_t)T * S);
        std::vector<uint32_t> offs(T + 1);
        for (uint32_t i = 0; i < T; i++) {
            offs[i] = i * S;
            for (uint32_t ti = 0; ti < S; ti++) {
                V3 p0 = vtx(ti, md.indices[3ull * i]), p1 = vtx(ti, md.indices[3ull * i + 1]),
                   p2 = vtx(ti, md.indices[3ull * i + 2]);
                bounds[(uint64_t)i * S + ti] = Box{vmin(p0, vmin(p1, p2)), vmax(p0, vmax(p1, p2))};
            }
        }
        offs[T] = T * S;
        BuildInput bi{T, bounds.data(), offs.data()};
        s.bvhs[m] = build_bvh4(bi, MESH_LEAF_TRIS, 0);
        const HostBvh4& bvh = s.bvhs[m];

        GpuMesh gm{};
        gm.root = append_bvh(s, bvh);
        gm.time_samples = S;
        gm.tri_stride = tri_record_bytes(S);
        gm.tri_offset = s.tris.size();
        // Triangle records in leaf order, vertices stored as (v0, v2, v1)
        // (triangle_mesh.rs:74-83).
        s.tris.resize(s.tris.size() + (uint64_t)gm.tri_stride * T, 0);
        for (uint32_t k = 0; k < T; k++) {
            uint32_t orig = bvh.order[k];
            uint8_t* rec = s.tris.data() + gm.tri_offset + (uint64_t)gm.tri_stride * k;
            float* f = reinterpret_cast<float*>(rec);
            const uint32_t slot[3] = {md.indices[3ull * orig], md.indices[3ull * orig + 2], md.indices[3ull * orig + 1]};
            for (uint32_t ti = 0; ti < S; ti++)
                for (int v = 0; v < 3; v++) {
                    V3 p = vtx(ti, slot[v]);
                    f[ti * 9 + v * 3 + 0] = p.x;
                    f[ti * 9 + v * 3 + 1] = p.y;
                    f[ti * 9 + v * 3 + 2] = p.z;
                }
            std::memcpy(rec + 36ull * S, &orig, 4);
        }
        s.meshes[m] = gm;
        s.max_depth = std::max(s.max_depth, bvh.depth);
        s.leaf_count += bvh.leaf_count;
        s.tri_count += T;
    }

    // ---- instances and the instance BVH (assembly.rs:288-395) ---------------
    std::vector<Box> ibounds;
    std::vector<uint32_t> ioffs{0};
    for (uint32_t i = 0; i < d->instance_count; i++) {
        const psy_instance_desc& id = d->instances[i];
        if (id.mesh >= d->mesh_count) { err = "instance " + std::to_string(i) + ": mesh out of range"; return PSY_E_INVALID; }
        if ((uint64_t)id.xform_start + id.xform_count > d->xform_count) {
            err = "instance " + std::to_string(i) + ": transform range out of bounds";
            return PSY_E_INVALID;
        }
        if (id.xform_count > MAX_TIME_SAMPLES) { err = "instance " + std::to_string(i) + ": too many transforms"; return PSY_E_RANGE; }
        std::vector<Xf> xf;
        for (uint32_t k = 0; k < id.xform_count; k++) xf.push_back(xf_at(d, id.xform_start + k));
        std::vector<Box> wb = transform_bounds(s.bvhs[id.mesh].root_bounds, xf.data(), id.xform_count);
        ibounds.insert(ibounds.end(), wb.begin(), wb.end());
        ioffs.push_back((uint32_t)ibounds.size());
    }
    BuildInput bi{d->instance_count, ibounds.data(), ioffs.data()};
    HostBvh4& top = s.bvhs[d->mesh_count];
    top = build_bvh4(bi, 1, 0);
    s.top_root = append_bvh(s, top);
    s.max_depth = std::max(s.max_depth, top.depth);
    s.leaf_count += top.leaf_count;

    for (uint32_t k = 0; k < d->instance_count; k++) {
        uint32_t orig = top.order[k];
        const psy_instance_desc& id = d->instances[orig];
        GpuInstance gi{};
        gi.mesh = id.mesh;
        gi.id = orig;
        gi.shader = id.shader;
        gi.xf_start = (uint32_t)s.xforms.size();
        gi.xf_count = id.xform_count;
        gi.inv_static = REF_EMPTY;
        for (uint32_t x = 0; x < id.xform_count; x++) s.xforms.push_back(to_gpu(xf_at(d, id.xform_start + x)));
        if (id.xform_count == 1) {
            // static_mat_space (triangle_mesh.rs:128-139)
            gi.inv_static = (uint32_t)s.inv_static.size();
            s.inv_static.push_back(to_gpu(xf_inverse(xf_at(d, id.xform_start))));
        }
        s.instances.push_back(gi);
        if (id.xform_count > 0) s.has_xf = true;
    }
    s.has_motion = !s.motion.empty();
    for (const GpuMesh& gm : s.meshes)
        if (gm.time_samples > 1) s.has_motion = true;

    // Pending-entry bound for the traversal stack: at most three siblings per
    // internal level along the current path, in both trees.
    uint32_t mesh_depth = 0;
    for (uint32_t m = 0; m < d->mesh_count; m++) mesh_depth = std::max(mesh_depth, s.bvhs[m].depth);
    s.max_stack = 3 * (top.depth + mesh_depth) + 2;
    s.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PSY_OK;
}

}  // namespace psy
