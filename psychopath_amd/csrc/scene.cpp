// Scene packing: TriangleMesh::from_verts_and_indices
// (src/surface/triangle_mesh.rs:29-116) and AssemblyBuilder::build
// (src/scene/assembly.rs:288-395) restated over the device layout.
#include "scene.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>

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

bool color_ok(const psy_color& c) {
    return c.kind <= PSY_COLOR_TEMPERATURE && std::isfinite(c.a) && std::isfinite(c.b) && std::isfinite(c.c);
}
GpuColor to_gpu(const psy_color& c) { return GpuColor{c.kind, c.a, c.b, c.c}; }

}  // namespace

std::vector<Xf> push_xforms(const std::vector<Xf>& top, const std::vector<Xf>& own) {
    if (own.empty()) return top;
    if (top.empty()) return own;
    std::vector<Xf> out;
    if (top.size() == own.size()) {
        for (size_t i = 0; i < own.size(); i++) out.push_back(xf_then(top[i], own[i]));
    } else if (top.size() > own.size()) {
        const float s = (float)(top.size() - 1);
        for (size_t i = 0; i < top.size(); i++)
            out.push_back(xf_then(top[i], lerp_slice_xf(own.data(), (uint32_t)own.size(), (float)i / s)));
    } else {
        const float s = (float)(own.size() - 1);
        for (size_t i = 0; i < own.size(); i++)
            out.push_back(xf_then(lerp_slice_xf(top.data(), (uint32_t)top.size(), (float)i / s), own[i]));
    }
    return out;
}

// Color::approximate_energy (src/color.rs:127-142), and the light objects'
// approximate_energy: the fold of their colours' estimates from 0 in list
// order, divided by the colour count (rectangle_light.rs:249-254,
// sphere_light.rs:198-203, distant_disk_light.rs:91-96).
float light_energy(const std::vector<psy_color>& cs) {
    float sum = 0.0f;
    for (const psy_color& c : cs) {
        float e;
        if (c.kind == PSY_COLOR_XYZ) e = c.b;
        else if (c.kind == PSY_COLOR_BLACKBODY) {
            float t2 = c.a * c.a;
            e = (t2 * t2) * c.b;
        } else e = c.b;
        sum = sum + e;
    }
    return sum / (float)cs.size();
}

// A light parameter's time samples: `n` values of `width` floats at `list`,
// or the one value at `one`.  Appends them to `samples` when there are
// several and returns the count stored there (0 for a single value, which
// is written to `one`).
uint32_t pack_samples(uint32_t n, const float* list, uint32_t width, float* one, std::vector<float>& samples) {
    if (n == 0) return 0;
    if (n == 1) {
        for (uint32_t k = 0; k < width; k++) one[k] = list[k];
        return 0;
    }
    samples.insert(samples.end(), list, list + (size_t)n * width);
    return n;
}

// A colour list: checks every colour and that they share a representation
// (Lerp for Color panics on mixed ones, color.rs:331).
bool colour_list(const psy_color& one, uint32_t n, const psy_color* cs, std::vector<psy_color>& out,
                 std::string& why) {
    out.clear();
    if (n == 0) out.push_back(one);
    else out.assign(cs, cs + n);
    for (const psy_color& c : out) {
        if (!color_ok(c)) {
            why = "bad colour";
            return false;
        }
        if (c.kind != out[0].kind) {
            why = "colour samples of different representations cannot be interpolated";
            return false;
        }
    }
    return true;
}

// Appends a checked colour list to `samples` (kind, a, b, c per colour) when
// it has several entries; returns the count stored (0 for one colour).
uint32_t pack_colours(const std::vector<psy_color>& cs, std::vector<float>& samples) {
    if (cs.size() < 2) return 0;
    for (const psy_color& c : cs) {
        samples.push_back((float)c.kind);
        samples.push_back(c.a);
        samples.push_back(c.b);
        samples.push_back(c.c);
    }
    return (uint32_t)cs.size();
}

namespace {

// merge_slices_append with union (src/algorithm.rs:233-262).
std::vector<Box> merge_union(const std::vector<Box>& a, const std::vector<Box>& b) {
    std::vector<Box> out;
    uint32_t na = (uint32_t)a.size(), nb = (uint32_t)b.size();
    if (na == nb) {
        for (uint32_t i = 0; i < na; i++) out.push_back(box_union(a[i], b[i]));
    } else if (na > nb) {
        float sc = (float)(na - 1);
        for (uint32_t i = 0; i < na; i++) out.push_back(box_union(a[i], lerp_slice_bx(b.data(), nb, (float)i / sc)));
    } else {
        float sc = (float)(nb - 1);
        for (uint32_t i = 0; i < nb; i++) out.push_back(box_union(lerp_slice_bx(a.data(), na, (float)i / sc), b[i]));
    }
    return out;
}

// Binary light tree (LightTree::from_objects, src/accel/light_tree.rs:
// 211-290): one light per leaf, SAH splits, bounds merged without the BVH's
// union rule, energies summed.
struct LtNode {
    bool leaf = false;
    std::vector<Box> bounds;
    float energy = 0.0f;
    uint32_t light = 0;
    std::unique_ptr<LtNode> kid[2];
};

std::unique_ptr<LtNode> lt_build(const BuildInput& in, const std::vector<float>& energy, uint32_t* keys, uint32_t n,
                                 uint32_t offset) {
    auto node = std::make_unique<LtNode>();
    if (n == 1) {
        node->leaf = true;
        node->bounds.assign(in.bounds + in.offsets[keys[0]], in.bounds + in.offsets[keys[0] + 1]);
        node->energy = energy[keys[0]];
        node->light = offset;
        return node;
    }
    uint32_t split = sah_split_keys(in, keys, n);
    node->kid[0] = lt_build(in, energy, keys, split, offset);
    node->kid[1] = lt_build(in, energy, keys + split, n - split, offset + split);
    node->bounds = merge_union(node->kid[0]->bounds, node->kid[1]->bounds);
    node->energy = node->kid[0]->energy + node->kid[1]->energy;
    return node;
}

// The same tree from a partition the device made (build_light_bvh2_device):
// node `id` of `t`, whose leaves hold the lights keys[a] at positions a.
std::unique_ptr<LtNode> lt_from_topology(const ArrayBvh2& t, uint32_t id, const BuildInput& in,
                                         const std::vector<float>& energy, const uint32_t* keys) {
    auto node = std::make_unique<LtNode>();
    if (t.meta[id] & ArrayBvh2::LEAF) {
        const uint32_t a = t.a[id];
        node->leaf = true;
        node->bounds.assign(in.bounds + in.offsets[keys[a]], in.bounds + in.offsets[keys[a] + 1]);
        node->energy = energy[keys[a]];
        node->light = a;
        return node;
    }
    node->kid[0] = lt_from_topology(t, t.a[id], in, energy, keys);
    node->kid[1] = lt_from_topology(t, t.b[id], in, energy, keys);
    node->bounds = merge_union(node->kid[0]->bounds, node->kid[1]->bounds);
    node->energy = node->kid[0]->energy + node->kid[1]->energy;
    return node;
}

// The collapse to arity 8 (light_tree.rs:292-372): a node's children are its
// descendants three levels down, or the leaves met before that, left to
// right.
void lt_collect(const LtNode* n, int level, std::vector<const LtNode*>& out) {
    if (level > 0 && !n->leaf) {
        lt_collect(n->kid[0].get(), level - 1, out);
        lt_collect(n->kid[1].get(), level - 1, out);
    } else {
        out.push_back(n);
    }
}

// Emits the collapsed tree of `n` with its root at light node `slot`;
// `leaf(light)` returns the light-leaf index of a leaf's light.
template <typename LeafFn>
void lt_emit(const LtNode* n, uint32_t slot, PackedScene& s, LeafFn&& leaf) {
    GpuLightNode g{};
    g.bounds_start = (uint32_t)s.light_boxes.size();
    g.bounds_len = (uint32_t)n->bounds.size();
    for (const Box& b : n->bounds) s.light_boxes.push_back(GpuBox{{b.lo.x, b.lo.y, b.lo.z}, {b.hi.x, b.hi.y, b.hi.z}});
    g.energy = n->energy;
    std::vector<const LtNode*> kids;
    if (n->leaf) {
        g.light = leaf(n->light);
    } else {
        lt_collect(n, 3, kids);
        g.child_first = (uint32_t)s.light_nodes.size();
        g.child_count = (uint32_t)kids.size();
        s.light_nodes.resize(s.light_nodes.size() + kids.size());
    }
    s.light_nodes[slot] = g;
    for (uint32_t i = 0; i < (uint32_t)kids.size(); i++) lt_emit(kids[i], g.child_first + i, s, leaf);
}

// ---- nested assemblies --------------------------------------------------
//
// A description with sub-assemblies is flattened into one single-level
// description: every assembly's meshes and lights (definitions placed depth
// first, the root's first) and, for every instance path, the mesh and light
// instances with the transforms the transform stack composes along it
// (tracer.rs:95-148).  The assemblies themselves stay for light selection:
// each instance path through an assembly is a group with the composed
// transforms into it and, per instance of the assembly, the flattened
// instance or sub-group it became.
constexpr uint32_t MAX_NESTING = 64;

struct DefInfo {
    uint32_t mesh_base = 0, rect_base = 0, sphere_base = 0;
    std::vector<Box> inst_bounds;   // per instance, its time samples in the assembly's space
    std::vector<uint32_t> inst_offs;
    std::vector<uint32_t> order;    // instance order after the assembly's instance BVH build
    std::vector<Box> bounds;        // that BVH's root bounds (Assembly::bounds, assembly.rs:144-148)
};
struct Group {
    const psy_scene_desc* def;
    std::vector<Xf> top;            // composed transforms into the assembly
    std::vector<uint32_t> target;   // per instance: flattened instance, or sub-group
    uint32_t depth;                 // 0 for the root assembly
};
struct Flat {
    std::vector<psy_mesh_desc> meshes;
    std::vector<psy_rect_light_desc> rects;
    std::vector<psy_sphere_light_desc> spheres;
    std::vector<psy_instance_desc> instances;
    std::vector<float> xforms;
    std::vector<const psy_scene_desc*> defs;  // depth first, the root first
    std::map<const psy_scene_desc*, uint32_t> def_index;
    std::vector<DefInfo> info;
    std::vector<Group> groups;                // [0]: the root assembly
    psy_scene_desc desc{};
};

int place_defs(const psy_scene_desc* d, Flat& f, uint32_t depth, std::string& err) {
    if (depth > MAX_NESTING) { err = "assemblies nest deeper than 64 levels"; return PSY_E_RANGE; }
    const std::string who = depth ? "sub-assembly: " : "";
    if (d->mesh_count > 0 && !d->meshes) { err = who + "meshes is NULL"; return PSY_E_INVALID; }
    if (d->instance_count > 0 && !d->instances) { err = who + "instances is NULL"; return PSY_E_INVALID; }
    if (d->xform_count > 0 && !d->xforms) { err = who + "xforms is NULL"; return PSY_E_INVALID; }
    if (d->instance_count > MAX_LEAF_FIRST) { err = who + "too many instances"; return PSY_E_RANGE; }
    if (d->rect_light_count > 0 && !d->rect_lights) { err = who + "rect_lights is NULL"; return PSY_E_INVALID; }
    if (d->sphere_light_count > 0 && !d->sphere_lights) { err = who + "sphere_lights is NULL"; return PSY_E_INVALID; }
    if (d->assembly_count > 0 && !d->assemblies) { err = who + "assemblies is NULL"; return PSY_E_INVALID; }
    if (f.def_index.count(d)) { err = "an assembly description is listed twice"; return PSY_E_INVALID; }
    const uint32_t id = (uint32_t)f.defs.size();
    f.defs.push_back(d);
    f.def_index[d] = id;
    f.info.emplace_back();
    f.info[id].mesh_base = (uint32_t)f.meshes.size();
    f.info[id].rect_base = (uint32_t)f.rects.size();
    f.info[id].sphere_base = (uint32_t)f.spheres.size();
    f.meshes.insert(f.meshes.end(), d->meshes, d->meshes + d->mesh_count);
    f.rects.insert(f.rects.end(), d->rect_lights, d->rect_lights + d->rect_light_count);
    f.spheres.insert(f.spheres.end(), d->sphere_lights, d->sphere_lights + d->sphere_light_count);
    for (uint32_t a = 0; a < d->assembly_count; a++) {
        const int rc = place_defs(&d->assemblies[a], f, depth + 1, err);
        if (rc) return rc;
    }
    return PSY_OK;
}

// Expands group `g`'s instances; sub-assembly instances open sub-groups.
int expand(uint32_t g, uint32_t depth, Flat& f, std::string& err) {
    const psy_scene_desc* d = f.groups[g].def;
    const uint32_t di = f.def_index.at(d);
    const std::vector<Xf> top = f.groups[g].top;
    std::vector<uint32_t> target(d->instance_count);
    for (uint32_t i = 0; i < d->instance_count; i++) {
        const psy_instance_desc& id = d->instances[i];
        const std::string who = (depth ? "sub-assembly instance " : "instance ") + std::to_string(i);
        if (id.kind > PSY_OBJECT_ASSEMBLY) { err = who + ": unknown object kind"; return PSY_E_INVALID; }
        const uint32_t limit = id.kind == PSY_OBJECT_MESH ? d->mesh_count
                               : id.kind == PSY_OBJECT_RECT_LIGHT ? d->rect_light_count
                               : id.kind == PSY_OBJECT_SPHERE_LIGHT ? d->sphere_light_count
                                                                     : d->assembly_count;
        if (id.mesh >= limit) { err = who + ": object out of range"; return PSY_E_INVALID; }
        if ((uint64_t)id.xform_start + id.xform_count > d->xform_count) {
            err = who + ": transform range out of bounds";
            return PSY_E_INVALID;
        }
        if (id.xform_count > MAX_TIME_SAMPLES) { err = who + ": too many transforms"; return PSY_E_RANGE; }
        std::vector<Xf> own;
        for (uint32_t k = 0; k < id.xform_count; k++) own.push_back(xf_at(d, id.xform_start + k));
        const std::vector<Xf> composed = push_xforms(top, own);
        if (composed.size() > MAX_TIME_SAMPLES) { err = who + ": too many composed transforms"; return PSY_E_RANGE; }
        if (id.kind == PSY_OBJECT_ASSEMBLY) {
            const uint32_t cg = (uint32_t)f.groups.size();
            f.groups.push_back(Group{&d->assemblies[id.mesh], composed, {}, depth + 1});
            target[i] = cg;
            const int rc = expand(cg, depth + 1, f, err);
            if (rc) return rc;
            continue;
        }
        if (id.kind != PSY_OBJECT_MESH && composed.empty()) {
            err = who + ": light instances need a transform";
            return PSY_E_INVALID;
        }
        psy_instance_desc out{};
        out.kind = id.kind;
        out.shader = id.shader;
        const DefInfo& info = f.info[di];
        out.mesh = id.mesh + (id.kind == PSY_OBJECT_MESH         ? info.mesh_base
                              : id.kind == PSY_OBJECT_RECT_LIGHT ? info.rect_base
                                                                 : info.sphere_base);
        out.xform_start = (uint32_t)(f.xforms.size() / 12);
        out.xform_count = (uint32_t)composed.size();
        for (const Xf& x : composed)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 4; c++) f.xforms.push_back(x.m[r][c]);
        target[i] = (uint32_t)f.instances.size();
        f.instances.push_back(out);
        if (f.instances.size() > MAX_LEAF_FIRST) { err = "too many instances"; return PSY_E_RANGE; }
    }
    f.groups[g].target = std::move(target);
    return PSY_OK;
}

// Object bounds of instance `id` of an assembly in the object's own space:
// a mesh's BVH root bounds, a light's bounds list (rectangle_light.rs:36-42,
// sphere_light.rs:35-40), or a sub-assembly's bounds.
std::vector<Box> object_bounds(const psy_scene_desc* d, const psy_instance_desc& id, const DefInfo& info,
                               const PackedScene& s, const Flat& f) {
    std::vector<Box> ob;
    if (id.kind == PSY_OBJECT_MESH) {
        ob = s.bvhs[info.mesh_base + id.mesh].root_bounds;
    } else if (id.kind == PSY_OBJECT_RECT_LIGHT) {
        const psy_rect_light_desc& l = d->rect_lights[id.mesh];
        for (uint32_t k = 0; k < std::max(1u, l.dims_count); k++) {
            const float w = l.dims_count ? l.dims[2 * k] : l.width, h = l.dims_count ? l.dims[2 * k + 1] : l.height;
            ob.push_back(Box{V3{w * -0.5f, h * -0.5f, 0.0f}, V3{w * 0.5f, h * 0.5f, 0.0f}});
        }
    } else if (id.kind == PSY_OBJECT_SPHERE_LIGHT) {
        const psy_sphere_light_desc& l = d->sphere_lights[id.mesh];
        for (uint32_t k = 0; k < std::max(1u, l.radius_count); k++) {
            const float r = l.radius_count ? l.radii[k] : l.radius;
            ob.push_back(Box{V3{-r, -r, -r}, V3{r, r, r}});
        }
    } else {
        ob = f.info[f.def_index.at(&d->assemblies[id.mesh])].bounds;
    }
    return ob;
}

float object_energy(const psy_scene_desc* d, const psy_instance_desc& id) {
    std::vector<psy_color> lc;
    std::string why;
    if (id.kind == PSY_OBJECT_RECT_LIGHT) {
        const psy_rect_light_desc& l = d->rect_lights[id.mesh];
        (void)colour_list(l.color, l.color_count, l.colors, lc, why);
    } else {
        const psy_sphere_light_desc& l = d->sphere_lights[id.mesh];
        (void)colour_list(l.color, l.color_count, l.colors, lc, why);
    }
    return light_energy(lc);
}

}  // namespace

// Packs a single-level description (the flattened scene `f`).
static int pack_flat(const psy_scene_desc* d, Flat& f, PackedScene& s, std::string& err, const BuildOptions& opt) {
    auto t0 = std::chrono::steady_clock::now();
    if (!d) { err = "scene description is NULL"; return PSY_E_INVALID; }
    if (d->mesh_count > 0 && !d->meshes) { err = "meshes is NULL"; return PSY_E_INVALID; }
    if (d->instance_count > 0 && !d->instances) { err = "instances is NULL"; return PSY_E_INVALID; }
    if (d->xform_count > 0 && !d->xforms) { err = "xforms is NULL"; return PSY_E_INVALID; }
    if (d->instance_count > MAX_LEAF_FIRST) { err = "too many instances"; return PSY_E_RANGE; }
    if (d->rect_light_count > 0 && !d->rect_lights) { err = "rect_lights is NULL"; return PSY_E_INVALID; }
    if (d->sphere_light_count > 0 && !d->sphere_lights) { err = "sphere_lights is NULL"; return PSY_E_INVALID; }
    auto size_ok = [](float v) { return v > 0.0f && std::isfinite(v); };
    std::vector<psy_color> cs;
    std::string why;
    for (uint32_t i = 0; i < d->rect_light_count; i++) {
        const psy_rect_light_desc& l = d->rect_lights[i];
        const std::string who = "rect light " + std::to_string(i) + ": ";
        if ((l.dims_count && !l.dims) || (l.color_count && !l.colors)) { err = who + "sample list is NULL"; return PSY_E_INVALID; }
        if (l.dims_count > MAX_TIME_SAMPLES || l.color_count > MAX_TIME_SAMPLES) {
            err = who + "too many time samples";
            return PSY_E_RANGE;
        }
        bool ok = l.dims_count ? true : size_ok(l.width) && size_ok(l.height);
        for (uint32_t k = 0; k < 2 * l.dims_count; k++) ok = ok && size_ok(l.dims[k]);
        if (!ok) { err = who + "bad size"; return PSY_E_INVALID; }
        if (!colour_list(l.color, l.color_count, l.colors, cs, why)) { err = who + why; return PSY_E_INVALID; }
        GpuRectLight g{l.width, l.height, (uint32_t)s.light_samples.size(), 0u, to_gpu(cs[0])};
        float wh[2] = {l.width, l.height};
        const uint32_t nd = pack_samples(l.dims_count, l.dims, 2, wh, s.light_samples);
        g.w = wh[0];
        g.h = wh[1];
        const uint32_t nc = pack_colours(cs, s.light_samples);
        g.counts = nd | (nc << 8);
        if (g.counts == 0) g.samples = 0;
        s.rect_lights.push_back(g);
    }
    for (uint32_t i = 0; i < d->sphere_light_count; i++) {
        const psy_sphere_light_desc& l = d->sphere_lights[i];
        const std::string who = "sphere light " + std::to_string(i) + ": ";
        if ((l.radius_count && !l.radii) || (l.color_count && !l.colors)) { err = who + "sample list is NULL"; return PSY_E_INVALID; }
        if (l.radius_count > MAX_TIME_SAMPLES || l.color_count > MAX_TIME_SAMPLES) {
            err = who + "too many time samples";
            return PSY_E_RANGE;
        }
        bool ok = l.radius_count ? true : size_ok(l.radius);
        for (uint32_t k = 0; k < l.radius_count; k++) ok = ok && size_ok(l.radii[k]);
        if (!ok) { err = who + "bad radius"; return PSY_E_INVALID; }
        if (!colour_list(l.color, l.color_count, l.colors, cs, why)) { err = who + why; return PSY_E_INVALID; }
        GpuSphereLight g{l.radius, (uint32_t)s.light_samples.size(), 0u, 0.0f, to_gpu(cs[0])};
        const uint32_t nr = pack_samples(l.radius_count, l.radii, 1, &g.radius, s.light_samples);
        const uint32_t nc = pack_colours(cs, s.light_samples);
        g.counts = nr | (nc << 8);
        if (g.counts == 0) g.samples = 0;
        s.sphere_lights.push_back(g);
    }

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
        if (md.normals) s.has_normals = true;
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
        bool built = false;
        if (opt.device && T >= opt.device_min_tris) {
            std::string why;
            double ms = 0;
            built = build_bvh4_device(bi, MESH_LEAF_TRIS, s.bvhs[m], why, &ms);
            if (built) {
                s.device_built++;
                s.device_build_ms += ms;
            }
        }
        if (!built) s.bvhs[m] = build_bvh4(bi, MESH_LEAF_TRIS, 0);
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
        gm.rec_map = (uint32_t)s.tri_rec.size();
        s.tri_rec.resize(s.tri_rec.size() + T);
        for (uint32_t k = 0; k < T; k++) s.tri_rec[gm.rec_map + bvh.order[k]] = k;
        gm.normal_base = PSY_NONE;
        gm.nidx_base = PSY_NONE;
        if (md.normals) {
            // Vertex-major normals and stored-order triples (triangle_mesh.rs:
            // 56-83); index limits keep both bases within 32 bits.
            if ((s.normals.size() / 3 + (uint64_t)V * S) >= PSY_NONE || (s.nidx.size() / 3 + T) >= PSY_NONE) {
                err = "mesh " + std::to_string(m) + ": too many vertices with normals";
                return PSY_E_RANGE;
            }
            gm.normal_base = (uint32_t)(s.normals.size() / 3);
            gm.nidx_base = (uint32_t)(s.nidx.size() / 3);
            for (uint32_t vi = 0; vi < V; vi++)
                for (uint32_t ti = 0; ti < S; ti++) {
                    const float* p = md.normals + 3ull * ((uint64_t)ti * V + vi);
                    s.normals.insert(s.normals.end(), p, p + 3);
                }
            for (uint32_t i = 0; i < T; i++) {
                s.nidx.push_back(md.indices[3ull * i]);
                s.nidx.push_back(md.indices[3ull * i + 2]);
                s.nidx.push_back(md.indices[3ull * i + 1]);
            }
        }
        s.meshes[m] = gm;
        s.max_depth = std::max(s.max_depth, bvh.depth);
        s.leaf_count += bvh.leaf_count;
        s.tri_count += T;
    }

    // Each assembly's instance bounds in its own space and its instance BVH
    // (AssemblyBuilder::build, assembly.rs:288-295), sub-assemblies (placed
    // after their parents) first.
    std::vector<HostBvh4> def_bvh(f.defs.size());
    for (uint32_t di = (uint32_t)f.defs.size(); di-- > 0;) {
        const psy_scene_desc* ad = f.defs[di];
        DefInfo& info = f.info[di];
        info.inst_bounds.clear();
        info.inst_offs.assign(1, 0u);
        for (uint32_t i = 0; i < ad->instance_count; i++) {
            const psy_instance_desc& id = ad->instances[i];
            std::vector<Xf> own;
            for (uint32_t k = 0; k < id.xform_count; k++) own.push_back(xf_at(ad, id.xform_start + k));
            const std::vector<Box> ob = object_bounds(ad, id, info, s, f);
            if (ob.empty()) {
                // The reference's builder lerps the instance's empty bounds
                // list and panics.
                err = "an instance of an assembly without instances";
                return PSY_E_INVALID;
            }
            const std::vector<Box> wb = transform_bounds(ob, own.data(), id.xform_count);
            info.inst_bounds.insert(info.inst_bounds.end(), wb.begin(), wb.end());
            info.inst_offs.push_back((uint32_t)info.inst_bounds.size());
        }
        BuildInput abi{ad->instance_count, info.inst_bounds.data(), info.inst_offs.data()};
        def_bvh[di] = build_bvh4(abi, 1, 0);
        info.order = def_bvh[di].order;
        info.bounds = def_bvh[di].root_bounds;
    }
    const bool nested = f.groups.size() > 1;
    uint32_t top_depth = 0;

    // ---- instances and the instance BVH (assembly.rs:288-395) ---------------
    if (!nested) {
    std::vector<Box> ibounds;
    std::vector<uint32_t> ioffs{0};
    for (uint32_t i = 0; i < d->instance_count; i++) {
        const psy_instance_desc& id = d->instances[i];
        const std::string who = "instance " + std::to_string(i);
        if (id.kind > PSY_OBJECT_SPHERE_LIGHT) { err = who + ": unknown object kind"; return PSY_E_INVALID; }
        const uint32_t limit = id.kind == PSY_OBJECT_MESH
                                   ? d->mesh_count
                                   : (id.kind == PSY_OBJECT_RECT_LIGHT ? d->rect_light_count : d->sphere_light_count);
        if (id.mesh >= limit) { err = who + ": object out of range"; return PSY_E_INVALID; }
        if ((uint64_t)id.xform_start + id.xform_count > d->xform_count) {
            err = who + ": transform range out of bounds";
            return PSY_E_INVALID;
        }
        if (id.xform_count > MAX_TIME_SAMPLES) { err = who + ": too many transforms"; return PSY_E_RANGE; }
        if (id.kind != PSY_OBJECT_MESH && id.xform_count == 0) {
            err = who + ": light instances need a transform";
            return PSY_E_INVALID;
        }
        std::vector<Xf> xf;
        for (uint32_t k = 0; k < id.xform_count; k++) xf.push_back(xf_at(d, id.xform_start + k));
        std::vector<Box> ob;
        if (id.kind == PSY_OBJECT_MESH) {
            ob = s.bvhs[id.mesh].root_bounds;
        } else if (id.kind == PSY_OBJECT_RECT_LIGHT) {
            // RectangleLight bounds, one box per dimensions sample
            // (rectangle_light.rs:36-42)
            const psy_rect_light_desc& l = d->rect_lights[id.mesh];
            for (uint32_t k = 0; k < std::max(1u, l.dims_count); k++) {
                const float w = l.dims_count ? l.dims[2 * k] : l.width, h = l.dims_count ? l.dims[2 * k + 1] : l.height;
                ob.push_back(Box{V3{w * -0.5f, h * -0.5f, 0.0f}, V3{w * 0.5f, h * 0.5f, 0.0f}});
            }
        } else {
            // SphereLight bounds, one box per radius sample
            // (sphere_light.rs:35-40)
            const psy_sphere_light_desc& l = d->sphere_lights[id.mesh];
            for (uint32_t k = 0; k < std::max(1u, l.radius_count); k++) {
                const float r = l.radius_count ? l.radii[k] : l.radius;
                ob.push_back(Box{V3{-r, -r, -r}, V3{r, r, r}});
            }
        }
        std::vector<Box> wb = transform_bounds(ob, xf.data(), id.xform_count);
        ibounds.insert(ibounds.end(), wb.begin(), wb.end());
        ioffs.push_back((uint32_t)ibounds.size());
    }
    BuildInput bi{d->instance_count, ibounds.data(), ioffs.data()};
    HostBvh4& top = s.bvhs[d->mesh_count];
    top = build_bvh4(bi, 1, 0);
    s.top_root = append_bvh(s, top);
    s.max_depth = std::max(s.max_depth, top.depth);
    s.leaf_count += top.leaf_count;

    s.inst_pos.resize(d->instance_count);
    for (uint32_t k = 0; k < d->instance_count; k++) {
        uint32_t orig = top.order[k];
        const psy_instance_desc& id = d->instances[orig];
        GpuInstance gi{};
        gi.mesh = id.mesh;
        gi.id = orig;
        gi.shader = id.shader;
        gi.kind = id.kind;
        s.inst_pos[orig] = k;
        if (id.kind != PSY_OBJECT_MESH) s.has_lights = true;
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
    top_depth = top.depth;
    } else {
        // Nested scene: every assembly keeps its own instance BVH, and every
        // instance path into an assembly (a group) gets that BVH's leaves as
        // instance records: the flattened object instances with their
        // composed transforms, and sub-assembly entries naming the path they
        // open (tracer.rs:87-150).
        const uint32_t nd = (uint32_t)f.defs.size();
        std::vector<uint32_t> def_root(nd);
        for (uint32_t di = 0; di < nd; di++) {
            def_root[di] = append_bvh(s, def_bvh[di]);
            s.leaf_count += def_bvh[di].leaf_count;
        }
        s.bvhs[d->mesh_count] = def_bvh[0];
        s.top_root = def_root[0];
        s.nested = true;
        s.has_xf = true;
        s.inst_pos.assign(d->instance_count, PSY_NONE);
        const uint32_t ng = (uint32_t)f.groups.size();
        s.groups.resize(ng);
        for (uint32_t g = 0; g < ng; g++) {
            const Group& G = f.groups[g];
            const uint32_t di = f.def_index.at(G.def);
            GpuGroup gg{def_root[di], (uint32_t)s.instances.size(), (uint32_t)s.xforms.size(), (uint32_t)G.top.size()};
            for (const Xf& x : G.top) s.xforms.push_back(to_gpu(x));
            s.groups[g] = gg;
            for (uint32_t k = 0; k < G.def->instance_count; k++) {
                const uint32_t orig = def_bvh[di].order[k];
                GpuInstance gi{};
                gi.inv_static = REF_EMPTY;
                if (G.def->instances[orig].kind == PSY_OBJECT_ASSEMBLY) {
                    gi.kind = PSY_OBJECT_ASSEMBLY;
                    gi.mesh = G.target[orig];
                    gi.id = PSY_NONE;
                    gi.shader = PSY_NONE;
                    s.instances.push_back(gi);
                    continue;
                }
                const uint32_t fi = G.target[orig];
                const psy_instance_desc& id = d->instances[fi];
                gi.mesh = id.mesh;
                gi.id = fi;
                gi.shader = id.shader;
                gi.kind = id.kind;
                s.inst_pos[fi] = (uint32_t)s.instances.size();
                if (id.kind != PSY_OBJECT_MESH) s.has_lights = true;
                gi.xf_start = (uint32_t)s.xforms.size();
                gi.xf_count = id.xform_count;
                for (uint32_t x = 0; x < id.xform_count; x++) s.xforms.push_back(to_gpu(xf_at(d, id.xform_start + x)));
                if (id.xform_count == 1) {
                    gi.inv_static = (uint32_t)s.inv_static.size();
                    s.inv_static.push_back(to_gpu(xf_inverse(xf_at(d, id.xform_start))));
                }
                s.instances.push_back(gi);
            }
        }
        if (s.instances.size() > MAX_LEAF_FIRST || ng > MAX_LEAF_FIRST) {
            err = "too many instance records in the nested scene";
            return PSY_E_RANGE;
        }
        for (uint32_t di = 0; di < nd; di++) s.max_depth = std::max(s.max_depth, def_bvh[di].depth);
    }
    // Light trees (AssemblyBuilder::build, assembly.rs:297-338): per group,
    // over its light instances and its sub-assembly instances whose trees
    // have energy, in the assembly's instance-BVH order; sub-groups (which
    // follow their parents) first, so their energies are known.
    const uint32_t ng = (uint32_t)f.groups.size();
    std::vector<std::unique_ptr<LtNode>> gtree(ng);
    std::vector<std::vector<uint32_t>> gents(ng);  // light order -> instance of the assembly
    std::vector<float> genergy(ng, 0.0f);
    for (uint32_t g = ng; g-- > 0;) {
        const Group& G = f.groups[g];
        const psy_scene_desc* ad = G.def;
        const DefInfo& info = f.info[f.def_index.at(ad)];
        std::vector<uint32_t> ents;
        std::vector<Box> lb;
        std::vector<uint32_t> lo{0};
        std::vector<float> le;
        for (uint32_t k : info.order) {
            const psy_instance_desc& id = ad->instances[k];
            float e;
            if (id.kind == PSY_OBJECT_MESH) continue;
            if (id.kind == PSY_OBJECT_ASSEMBLY) {
                e = genergy[G.target[k]];
                if (!(e > 0.0f)) continue;
            } else {
                e = object_energy(ad, id);
            }
            ents.push_back(k);
            lb.insert(lb.end(), info.inst_bounds.begin() + info.inst_offs[k], info.inst_bounds.begin() + info.inst_offs[k + 1]);
            lo.push_back((uint32_t)lb.size());
            le.push_back(e);
        }
        if (ents.empty()) continue;
        if (G.depth >= MAX_LIGHT_LEVELS) {
            err = "lights nested in more than " + std::to_string(MAX_LIGHT_LEVELS - 1) + " levels of sub-assemblies";
            return PSY_E_RANGE;
        }
        std::vector<uint32_t> keys(ents.size());
        for (uint32_t i = 0; i < (uint32_t)keys.size(); i++) keys[i] = i;
        BuildInput lin{(uint32_t)keys.size(), lb.data(), lo.data()};
        bool built = false;
        if (opt.device && keys.size() >= std::max(2u, opt.device_min_tris)) {
            ArrayBvh2 topo;
            std::vector<uint32_t> order;
            std::string why;
            if (build_light_bvh2_device(lin, topo, order, why)) {
                keys = order;
                gtree[g] = lt_from_topology(topo, 0, lin, le, keys.data());
                s.device_light_trees++;
                built = true;
            }
        }
        if (!built) gtree[g] = lt_build(lin, le, keys.data(), (uint32_t)keys.size(), 0);
        genergy[g] = gtree[g]->energy;
        gents[g].resize(keys.size());
        for (uint32_t i = 0; i < (uint32_t)keys.size(); i++) gents[g][i] = ents[keys[i]];
    }
    if (gtree[0]) {
        auto push_list = [&](const std::vector<Xf>& l, uint32_t& start, uint32_t& count) {
            start = (uint32_t)s.light_xforms.size();
            count = (uint32_t)l.size();
            for (const Xf& x : l) s.light_xforms.push_back(to_gpu(x));
        };
        std::deque<std::pair<uint32_t, uint32_t>> pending;  // (leaf, sub-group) whose tree is not emitted yet
        auto emit_group = [&](uint32_t g, uint32_t slot) {
            const Group& G = f.groups[g];
            const psy_scene_desc* ad = G.def;
            lt_emit(gtree[g].get(), slot, s, [&](uint32_t light) {
                const uint32_t k = gents[g][light];
                const psy_instance_desc& id = ad->instances[k];
                GpuLightLeaf lf{};
                const uint32_t leaf = (uint32_t)s.light_leaves.size();
                if (id.kind == PSY_OBJECT_ASSEMBLY) {
                    lf.kind = LIGHT_LEAF_ASSEMBLY;
                    push_list(f.groups[G.target[k]].top, lf.top_start, lf.top_count);
                    pending.emplace_back(leaf, G.target[k]);
                } else {
                    lf.kind = LIGHT_LEAF_LIGHT;
                    lf.target = s.inst_pos[G.target[k]];
                    push_list(G.top, lf.top_start, lf.top_count);
                    std::vector<Xf> own;
                    for (uint32_t x = 0; x < id.xform_count; x++) own.push_back(xf_at(ad, id.xform_start + x));
                    push_list(own, lf.own_start, lf.own_count);
                }
                s.light_leaves.push_back(lf);
                return leaf;
            });
        };
        s.light_nodes.resize(1);
        emit_group(0, 0);
        while (!pending.empty()) {
            const auto [leaf, g] = pending.front();
            pending.pop_front();
            const uint32_t slot = (uint32_t)s.light_nodes.size();
            s.light_nodes.resize(slot + 1);
            emit_group(g, slot);
            s.light_leaves[leaf].target = slot;
        }
    }

    s.has_motion = !s.motion.empty();
    for (const GpuMesh& gm : s.meshes)
        if (gm.time_samples > 1) s.has_motion = true;

    // Pending-entry bound for the traversal stack: at most three siblings per
    // internal level along the current path, in both trees.
    uint32_t mesh_depth = 0;
    for (uint32_t m = 0; m < d->mesh_count; m++) mesh_depth = std::max(mesh_depth, s.bvhs[m].depth);
    if (!nested) {
        s.max_stack = 3 * (top_depth + mesh_depth) + 2;
    } else {
        // Along every path: the assemblies' BVHs, a level marker per
        // sub-assembly entered, and the deepest mesh BVH at its end.
        std::vector<uint32_t> need(f.groups.size(), 0);
        for (uint32_t g = (uint32_t)f.groups.size(); g-- > 0;) {
            const Group& G = f.groups[g];
            uint32_t below = 0;
            for (uint32_t k = 0; k < G.def->instance_count; k++) {
                const psy_instance_desc& id = G.def->instances[k];
                if (id.kind == PSY_OBJECT_ASSEMBLY) below = std::max(below, 1 + need[G.target[k]]);
                else if (id.kind == PSY_OBJECT_MESH) below = std::max(below, 3 * s.bvhs[d->instances[G.target[k]].mesh].depth);
            }
            need[g] = 3 * def_bvh[f.def_index.at(G.def)].depth + below;
        }
        s.max_stack = need[0] + 2;
    }

    // Device form of the node array: each node's traversal code also rides
    // in bits 28..30 of its first two child references (layout.h), and the
    // order table it indexes travels with the scene.
    if (s.nodes.size() > MAX_NODES) {
        err = "the scene's BVHs have more than 2^28 nodes";
        return PSY_E_RANGE;
    }
    for (GpuNode& n : s.nodes) {
        const uint32_t code = (n.meta >> 16) & 0x3fu;
        n.child[0] |= (code & 7u) << REF_CODE_SHIFT;
        n.child[1] |= (code >> 3) << REF_CODE_SHIFT;
    }
    s.order_table.assign(traversal_order_table(), traversal_order_table() + 8 * 48);
    s.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PSY_OK;
}

int pack_scene(const psy_scene_desc* d, PackedScene& s, std::string& err, const BuildOptions& opt) {
    if (!d) { err = "scene description is NULL"; return PSY_E_INVALID; }
    Flat f;
    int rc = place_defs(d, f, 0, err);
    if (rc) return rc;
    f.groups.push_back(Group{d, {}, {}, 0});
    rc = expand(0, 0, f, err);
    if (rc) return rc;
    f.desc.mesh_count = (uint32_t)f.meshes.size();
    f.desc.meshes = f.meshes.data();
    f.desc.instance_count = (uint32_t)f.instances.size();
    f.desc.instances = f.instances.data();
    f.desc.xform_count = (uint32_t)(f.xforms.size() / 12);
    f.desc.xforms = f.xforms.data();
    f.desc.rect_light_count = (uint32_t)f.rects.size();
    f.desc.rect_lights = f.rects.data();
    f.desc.sphere_light_count = (uint32_t)f.spheres.size();
    f.desc.sphere_lights = f.spheres.data();
    return pack_flat(&f.desc, f, s, err, opt);
}

}  // namespace psy
