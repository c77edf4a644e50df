// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
#include "psyref_render.h"

#include <cstdio>
#include <cstdlib>

#include <atomic>
#include <chrono>
#include <thread>

namespace psyref {

const float RectangleLight::SIMPLE_SAMPLING_THRESHOLD = 0.01f;  // rectangle_light.rs:21

const Halton& halton() {
    static Halton h;
    return h;
}

// ---- light_tree.rs:211-372 ---------------------------------------------------
namespace {
struct LTBuilder {
    struct BNode {
        bool leaf;
        size_t b0, b1;
        float energy;
        size_t child_index;
    };
    std::vector<BNode> nodes;
    std::vector<BBox> bounds;
    size_t depth = 0;
    const std::function<BoundsSlice(uint32_t)>* bounds_of;
    const std::function<float(uint32_t)>* energy_of;

    std::pair<size_t, std::pair<size_t, size_t>> build(size_t offset, size_t d, uint32_t* objs, size_t n) {
        size_t me = nodes.size();
        if (n == 0) return {0, {0, 0}};
        if (n == 1) {
            size_t bi = bounds.size();
            BoundsSlice bs = (*bounds_of)(objs[0]);
            bounds.insert(bounds.end(), bs.ptr, bs.ptr + bs.len);
            nodes.push_back(BNode{true, bi, bounds.size(), (*energy_of)(objs[0]), offset});
            if (depth < d) depth = d;
            return {me, {bi, bounds.size()}};
        }
        nodes.push_back(BNode{false, 0, 0, 0.0f, 0});
        auto split = sah_split(objs, n, [&](uint32_t k) { return (*bounds_of)(k); });
        size_t s = split.first;
        auto c1 = build(offset, d + 1, objs, s);
        auto c2 = build(offset + s, d + 1, objs + s, n - s);
        size_t bi = bounds.size();
        // merge_slices_append with union (algorithm.rs:233-262)
        std::vector<BBox> merged;
        const BBox* a = bounds.data() + c1.second.first;
        size_t na = c1.second.second - c1.second.first;
        const BBox* b = bounds.data() + c2.second.first;
        size_t nb = c2.second.second - c2.second.first;
        if (na == nb) {
            for (size_t i = 0; i < na; i++) merged.push_back(bunion(a[i], b[i]));
        } else if (na > nb) {
            float sc = (float)(na - 1);
            for (size_t i = 0; i < na; i++) merged.push_back(bunion(a[i], lerp_slice(b, nb, (float)i / sc)));
        } else {
            float sc = (float)(nb - 1);
            for (size_t i = 0; i < nb; i++) merged.push_back(bunion(lerp_slice(a, na, (float)i / sc), b[i]));
        }
        bounds.insert(bounds.end(), merged.begin(), merged.end());
        float energy = nodes[me + 1].energy + nodes[c2.first].energy;
        nodes[me] = BNode{false, bi, bounds.size(), energy, c2.first};
        (void)c1;
        return {me, {bi, bounds.size()}};
    }

    static const size_t ARITY_LOG2 = 3;
    size_t child_count(size_t level, size_t idx) const {
        if (level > 0) {
            if (nodes[idx].leaf) return 1;
            return child_count(level - 1, idx + 1) + child_count(level - 1, nodes[idx].child_index);
        }
        return 1;
    }
    std::pair<size_t, size_t> nth_child(size_t level, size_t idx, size_t n) const {
        if (level > 0 && !nodes[idx].leaf) {
            auto r = nth_child(level - 1, idx + 1, n);
            if (r.second == 0) return {r.first, 0};
            return nth_child(level - 1, nodes[idx].child_index, r.second - 1);
        }
        return {idx, n};
    }
    void construct(size_t idx, LightTree::Node& out) const {
        const BNode& b = nodes[idx];
        out.bounds.assign(bounds.begin() + b.b0, bounds.begin() + b.b1);
        out.energy = b.energy;
        if (b.leaf) {
            out.leaf = true;
            out.light_index = b.child_index;
            return;
        }
        out.leaf = false;
        size_t cc = child_count(ARITY_LOG2, idx);
        out.children.resize(cc);
        for (size_t i = 0; i < cc; i++) construct(nth_child(ARITY_LOG2, idx, i).first, out.children[i]);
    }
};
}  // namespace

void LightTree::build(uint32_t* objects, size_t n, const std::function<BoundsSlice(uint32_t)>& bounds_of,
                      const std::function<float(uint32_t)>& energy_of) {
    has_root = false;
    if (n == 0) return;
    LTBuilder b;
    b.bounds_of = &bounds_of;
    b.energy_of = &energy_of;
    b.build(0, 0, objects, n);
    b.construct(0, root);
    has_root = true;
    depth = b.depth;
}

bool LightTree::select(Vec3 inc, Vec3 pos, Vec3 nor, Vec3 nor_g, const Closure& sc, float time, float n,
                       size_t& index, float& prob, float& whittled_n) const {
    if (!has_root) return false;
    auto node_prob = [&](const Node& node) {
        BBox bbox = lerp_slice(node.bounds.data(), node.bounds.size(), time);
        Vec3 d = bbox.center() - pos;
        float r2 = length2(bbox.max - bbox.min) * 0.25f;
        float inv_surface_area = 1.0f / r2;
        float approx = closure_estimate_sphere_light(sc, inc, d, r2, nor, nor_g);
        return (node.energy * inv_surface_area) * approx;
    };
    const Node* node = &root;
    float tot_prob = 1.0f;
    while (!node->leaf) {
        const size_t len = node->children.size();
        float ps[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        float total = 0.0f;
        for (size_t i = 0; i < len; i++) {
            float p = node_prob(node->children[i]);
            ps[i] = p;
            total += p;
        }
        if (total <= 0.0f) {
            float p = 1.0f / (float)len;
            for (float& q : ps) q = p;
        } else {
            for (float& q : ps) q /= total;
        }
        float base = 0.0f;
        for (size_t i = 0; i < 8; i++) {
            float p = ps[i];
            if ((n <= base + p) || (i == len - 1)) {
                tot_prob *= p;
                node = &node->children[i];
                n = (n - base) / p;
                break;
            }
            base += p;
        }
    }
    index = node->light_index;
    prob = tot_prob;
    whittled_n = n;
    return true;
}

// AssemblyBuilder::build, light part (assembly.rs:297-338), for every
// assembly, sub-assemblies first: the light instances are the instances of
// light objects and of sub-assemblies whose own light tree has energy, the
// latter weighted by that energy.
namespace {
void build_assembly_lights(const Assembly* a, std::map<const Assembly*, AssemblyLights>& out) {
    for (const auto& sub : a->assemblies) build_assembly_lights(sub.get(), out);
    AssemblyLights& al = out[a];
    al.instances.clear();
    auto energy_of = [&](const Instance& inst) {
        if (inst.is_assembly) return out.at(a->assemblies[inst.data_index].get()).tree.approximate_energy();
        return a->light_energy(inst);
    };
    for (const Instance& inst : a->instances) {
        if (inst.is_assembly) {
            if (out.at(a->assemblies[inst.data_index].get()).tree.approximate_energy() > 0.0f)
                al.instances.push_back(inst);
        } else if (inst.obj_kind != OBJ_MESH) {
            al.instances.push_back(inst);
        }
    }
    std::vector<uint32_t> keys(al.instances.size());
    for (size_t i = 0; i < keys.size(); i++) keys[i] = (uint32_t)i;
    const std::vector<Instance> li = al.instances;
    al.tree.build(
        keys.data(), keys.size(),
        [&](uint32_t k) {
            size_t id = li[k].id;
            return BoundsSlice{a->inst_bounds.data() + a->inst_bound_start[id],
                               a->inst_bound_start[id + 1] - a->inst_bound_start[id]};
        },
        [&](uint32_t k) { return energy_of(li[k]); });
    for (size_t i = 0; i < keys.size(); i++) al.instances[i] = li[keys[i]];
}
}  // namespace

void RenderScene::build_lights() {
    lights.clear();
    build_assembly_lights(root, lights);
}

// ---- math.rs:67-132 ----------------------------------------------------------
float probit(float n_f, float width) {
    double n = (double)n_f;
    const double A1 = -3.969683028665376e+01, A2 = 2.209460984245205e+02, A3 = -2.759285104469687e+02,
                 A4 = 1.383577518672690e+02, A5 = -3.066479806614716e+01, A6 = 2.506628277459239e+00;
    const double B1 = -5.447609879822406e+01, B2 = 1.615858368580409e+02, B3 = -1.556989798598866e+02,
                 B4 = 6.680131188771972e+01, B5 = -1.328068155288572e+01;
    const double C1 = -7.784894002430293e-03, C2 = -3.223964580411365e-01, C3 = -2.400758277161838e+00,
                 C4 = -2.549732539343734e+00, C5 = 4.374664141464968e+00, C6 = 2.938163982698783e+00;
    const double D1 = 7.784695709041462e-03, D2 = 3.224671290700398e-01, D3 = 2.445134137142996e+00,
                 D4 = 3.754408661907416e+00;
    const double N_LOW = 0.02425, N_HIGH = 1.0 - N_LOW;
    double x;
    if (0.0 < n && n < N_LOW) {
        double q = std::sqrt(-2.0 * std::log(n));
        x = (((((C1 * q + C2) * q + C3) * q + C4) * q + C5) * q + C6) / ((((D1 * q + D2) * q + D3) * q + D4) * q + 1.0);
    } else if (n <= N_HIGH) {
        double q = n - 0.5;
        double r = q * q;
        x = (((((A1 * r + A2) * r + A3) * r + A4) * r + A5) * r + A6) * q /
            (((((B1 * r + B2) * r + B3) * r + B4) * r + B5) * r + 1.0);
    } else if (n < 1.0) {
        double q = std::sqrt(-2.0 * std::log(1.0 - n));
        x = -(((((C1 * q + C2) * q + C3) * q + C4) * q + C5) * q + C6) /
            ((((D1 * q + D2) * q + D3) * q + D4) * q + 1.0);
    } else if (n == 0.0) {
        x = -INFINITY;
    } else if (n == 1.0) {
        x = INFINITY;
    } else {
        x = NAN;
    }
    return (float)x * width;
}

namespace {

// Scene::sample_lights + Assembly::sample_lights (scene/mod.rs:48-121,
// scene/assembly.rs:46-141) for a single-level root assembly.
enum LightSampleKind { LS_NONE, LS_DISTANT, LS_SURFACE };
struct SceneLightSample {
    LightSampleKind kind = LS_NONE;
    Vec4 color;
    Vec3 direction;                  // distant
    Vec3 point, normal;              // surface
    float point_err = 0;
    float pdf = 0, selection_pdf = 0;
};

float weighted_choice_energy(const std::vector<DistantDiskLight>& ls, float n, size_t& idx) {
    float total = 0.0f;
    for (const auto& l : ls) total = total + l.approximate_energy();
    float nn = n * total;
    float x = 0.0f;
    for (size_t i = 0; i < ls.size(); i++) {
        float w = ls[i].approximate_energy();
        x += w;
        if (x > nn) {
            idx = i;
            return w / total;
        }
    }
    idx = ls.size() - 1;  // unreachable!() in the reference
    return ls.back().approximate_energy() / total;
}

// Assembly::sample_lights (assembly.rs:46-141): selects in this assembly's
// light tree with the shading data taken into its space (the lerped top of
// the transform stack), then samples the chosen light under the stack's
// transform times the instance's, or recurses into the chosen sub-assembly
// with its instance transforms pushed.  Returns false when nothing is
// selected; `sel_pdf` is the product of the selection probabilities.
bool assembly_sample_lights(const RenderScene& sc, const Assembly* a, TransformStack& xs, float n, float u, float v,
                            float wavelength, float time, const Isect& is, const Closure& closure, LightSample& s,
                            float& sel_pdf) {
    size_t ntop;
    const Transform* top = xs.top(ntop);
    const Transform sel_xform = ntop ? lerp_slice(top, ntop, time) : Transform();
    const AssemblyLights& al = sc.lights_of(a);
    size_t li;
    float p, whittled;
    if (!al.tree.select(xform_vector(is.incoming, sel_xform), xform_point(is.pos, sel_xform),
                        xform_normal(is.nor, sel_xform), xform_normal(is.nor_g, sel_xform), closure, time, n, li, p,
                        whittled))
        return false;
    const Instance& inst = al.instances[li];
    if (!inst.is_assembly) {
        Transform xform;
        if (inst.has_xforms) {
            const Transform x = lerp_slice(a->xforms.data() + inst.xf_start, inst.xf_end - inst.xf_start, time);
            xform = ntop ? xform_compose(lerp_slice(top, ntop, time), x) : x;
        } else if (ntop) {
            xform = lerp_slice(top, ntop, time);
        }
        s = inst.obj_kind == OBJ_RECT_LIGHT
                ? a->rect_lights[inst.data_index].at(time).sample_from_point(xform, is.pos, u, v, wavelength)
                : a->sphere_lights[inst.data_index].at(time).sample_from_point(xform, is.pos, u, v, wavelength);
        sel_pdf = p;
        return true;
    }
    if (inst.has_xforms) xs.push(a->xforms.data() + inst.xf_start, inst.xf_end - inst.xf_start);
    float sub_pdf = 0.0f;
    const bool ok = assembly_sample_lights(sc, a->assemblies[inst.data_index].get(), xs, whittled, u, v, wavelength,
                                           time, is, closure, s, sub_pdf);
    if (inst.has_xforms) xs.pop();
    sel_pdf = sub_pdf * p;
    return ok;
}

SceneLightSample sample_lights(const RenderScene& sc, float n, float u, float v, float wavelength, float time,
                               const Isect& is, const Closure& closure) {
    SceneLightSample out;
    float wl_sum = 0.0f;
    for (const auto& l : sc.world.lights) wl_sum = wl_sum + l.approximate_energy();
    float wl_energy = wl_sum <= 0.0f ? 0.0f : 1.0f;
    float ll_energy = sc.lights_of(sc.root).tree.approximate_energy() <= 0.0f ? 0.0f : 1.0f;
    float tot = wl_energy + ll_energy;
    if (tot <= 0.0f) return out;
    float wl_prob = wl_energy / tot;
    if (n < wl_prob) {
        float nn = n / wl_prob;
        size_t i;
        float p = weighted_choice_energy(sc.world.lights, nn, i);
        Vec4 col;
        Vec3 dir;
        float pdf;
        sc.world.lights[i].at(time).sample_from_point(u, v, wavelength, col, dir, pdf);
        out.kind = LS_DISTANT;
        out.color = col;
        out.direction = dir;
        out.pdf = pdf;
        out.selection_pdf = p * wl_prob;
        return out;
    }
    float nn = (n - wl_prob) / (1.0f - wl_prob);
    // Assembly::sample_lights from the root with an empty transform stack.
    TransformStack xs;
    LightSample s;
    float sel_pdf = 0.0f;
    if (!assembly_sample_lights(sc, sc.root, xs, nn, u, v, wavelength, time, is, closure, s, sel_pdf)) return out;
    out.kind = LS_SURFACE;
    out.color = s.color;
    out.point = s.point;
    out.normal = s.normal;
    out.point_err = s.point_err;
    out.pdf = s.pdf;
    out.selection_pdf = sel_pdf * (1.0f - wl_prob);
    return out;
}

// renderer.rs:362-447
enum PathEvent { EV_CAMERA, EV_BOUNCE, EV_SHADOW };
struct LightPath {
    PathEvent event = EV_CAMERA;
    uint32_t bounce_count = 0;
    uint32_t sampling_seed = 0;
    uint32_t px = 0, py = 0, sample_number = 0, dim_offset = 0;
    float time = 0, wavelength = 0;
    bool has_next_bounce = false;
    Ray next_bounce_ray{};
    Vec4 next_attenuation_fac = Vec4::splat(1.0f);
    float closure_sample_pdf = 1.0f;
    Vec4 light_attenuation = Vec4::splat(1.0f);
    Vec4 pending_color_addition;
    Vec4 color;

    void next_lds_sequence() {
        dim_offset = 0;
        sampling_seed += 1;
    }
    void next_lds_samp(float out[4]) { get_sample_4d(sample_number, dim_offset++, px, py, sampling_seed, out); }
};

// The colour the reference gives mesh surfaces without a shader:
// Emit(rec709_to_xyz((1, 0, 1))) (tracer.rs:161-164).  rec709_to_xyz's
// constants come from the colour crate's build script (Rec.709 primaries,
// D65 white, RGBtoXYZ of the ACES CTL, printed to ten decimals and read as
// f32 literals; sub_crates/color/build.rs:14-24, 78-99, 178-225).
static Color unassigned_color() {
    const double rx = 0.640, ry = 0.330, gx = 0.300, gy = 0.600, bx = 0.150, by = 0.060;
    const double wx = 0.3127, wy = 0.3290, Y = 1.0;
    const double X = wx * Y / wy, Z = (1.0 - wx - wy) * Y / wy;
    const double d = rx * (by - gy) + bx * (gy - ry) + gx * (ry - by);
    const double sr = (X * (by - gy) - gx * (Y * (by - 1.0) + by * (X + Z)) + bx * (Y * (gy - 1.0) + gy * (X + Z))) / d;
    const double sg = (X * (ry - by) + rx * (Y * (by - 1.0) + by * (X + Z)) - bx * (Y * (ry - 1.0) + ry * (X + Z))) / d;
    const double sb = (X * (gy - ry) - rx * (Y * (gy - 1.0) + gy * (X + Z)) + gx * (Y * (ry - 1.0) + ry * (X + Z))) / d;
    const double m[3][3] = {{sr * rx, sg * gx, sb * bx},
                            {sr * ry, sg * gy, sb * by},
                            {sr * (1.0 - rx - ry), sg * (1.0 - gx - gy), sb * (1.0 - bx - by)}};
    float f[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            char buf[64];
            snprintf(buf, sizeof buf, "%.10f", m[i][j]);
            f[i][j] = strtof(buf, nullptr);
        }
    const float r = 1.0f, g = 0.0f, b = 1.0f;
    Color c;
    c.kind = COLOR_XYZ;
    c.a = (r * f[0][0] + g * f[0][1]) + b * f[0][2];
    c.b = (r * f[1][0] + g * f[1][1]) + b * f[1][2];
    c.c = (r * f[2][0] + g * f[2][1]) + b * f[2][2];
    return c;
}

Closure closure_of(const RenderScene& sc, const Isect& is) {
    if (is.light) {
        Closure c;
        c.kind = CLOSURE_EMIT;
        c.color = is.emit;
        return c;
    }
    if (is.shader < 0) {
        static const Color magenta = unassigned_color();
        Closure c;
        c.kind = CLOSURE_EMIT;
        c.color = magenta;
        return c;
    }
    return sc.shaders[(size_t)is.shader];
}

// LightPath::next (renderer.rs:449-692)
bool path_next(LightPath& p, const RenderScene& sc, const Isect& is, RayBatch& rays, size_t ri, uint32_t max_bounces) {
    if (p.event == EV_CAMERA || p.event == EV_BOUNCE) {
        if (is.kind == ISECT_HIT) {
            Closure closure = closure_of(sc, is);
            if (closure.kind == CLOSURE_EMIT) {
                Vec4 c = closure.color.to_spectral_sample(p.wavelength);
                if (p.event == EV_CAMERA) {
                    p.color = p.color + c;
                } else {
                    float mis_pdf = power_heuristic(p.closure_sample_pdf, is.sample_pdf);
                    p.color = p.color + ((c * p.light_attenuation) / mis_pdf);
                }
                return false;
            }
            p.light_attenuation = p.light_attenuation / p.closure_sample_pdf;

            // Light ray
            p.next_lds_sequence();
            float ls4[4];
            p.next_lds_samp(ls4);
            SceneLightSample li = sample_lights(sc, ls4[0], ls4[1], ls4[2], p.wavelength, p.time, is, closure);
            bool found_light;
            if (li.kind == LS_NONE || li.pdf <= 0.0f || li.selection_pdf <= 0.0f) {
                found_light = false;
            } else {
                Vec4 att;
                float closure_pdf;
                Ray shadow{};
                if (li.kind == LS_DISTANT) {
                    closure_evaluate(closure, rays.dir(ri), li.direction, is.nor, is.nor_g, p.wavelength, att,
                                     closure_pdf);
                    shadow.orig = robust_ray_origin(is.pos, is.pos_err, normalized(is.nor_g), li.direction);
                    shadow.dir = li.direction;
                    shadow.time = p.time;
                    shadow.wavelength = p.wavelength;
                    shadow.max_t = INFINITY;
                } else {
                    Vec3 dir = li.point - is.pos;
                    closure_evaluate(closure, rays.dir(ri), dir, is.nor, is.nor_g, p.wavelength, att, closure_pdf);
                    Vec3 offset_pos = robust_ray_origin(is.pos, is.pos_err, normalized(is.nor_g), dir);
                    Vec3 offset_end = robust_ray_origin(li.point, li.point_err, normalized(li.normal), neg(dir));
                    shadow.orig = offset_pos;
                    shadow.dir = offset_end - offset_pos;
                    shadow.time = p.time;
                    shadow.wavelength = p.wavelength;
                    shadow.max_t = 1.0f;
                }
                if (max_element4(att) <= 0.0f) {
                    found_light = false;
                } else {
                    float light_mis_pdf = power_heuristic(li.pdf, closure_pdf);
                    p.pending_color_addition =
                        ((li.color * att) * p.light_attenuation) / (light_mis_pdf * li.selection_pdf);
                    rays.set_from_ray(shadow, true, ri);
                    found_light = true;
                }
            }

            // Bounce ray
            bool do_bounce;
            if (p.bounce_count < max_bounces) {
                p.bounce_count += 1;
                p.next_lds_sequence();
                float b4[4];
                p.next_lds_samp(b4);
                Vec3 dir;
                Vec4 filter;
                float pdf;
                closure_sample(closure, is.incoming, is.nor, is.nor_g, b4[0], b4[1], b4[2], p.wavelength, dir, filter,
                               pdf);
                if ((pdf > 0.0f) && (max_element4(filter) > 0.0f)) {
                    p.next_attenuation_fac = filter;
                    p.closure_sample_pdf = pdf;
                    Vec3 offset_pos = robust_ray_origin(is.pos, is.pos_err, normalized(is.nor_g), dir);
                    p.next_bounce_ray = Ray{offset_pos, dir, p.time, p.wavelength, INFINITY};
                    p.has_next_bounce = true;
                    do_bounce = true;
                } else {
                    do_bounce = false;
                }
            } else {
                p.has_next_bounce = false;
                do_bounce = false;
            }

            if (found_light) {
                p.event = EV_SHADOW;
                return true;
            } else if (do_bounce) {
                rays.set_from_ray(p.next_bounce_ray, false, ri);
                p.event = EV_BOUNCE;
                p.light_attenuation = p.light_attenuation * p.next_attenuation_fac;
                return true;
            }
            return false;
        }
        Vec4 bg = sc.world.background.to_spectral_sample(p.wavelength);
        p.color = p.color + ((bg * p.light_attenuation) / p.closure_sample_pdf);
        return false;
    }
    // Shadow ray (renderer.rs:670-690)
    if (is.kind == ISECT_MISS) p.color = p.color + p.pending_color_addition;
    if (p.has_next_bounce) {
        rays.set_from_ray(p.next_bounce_ray, false, ri);
        p.light_attenuation = p.light_attenuation * p.next_attenuation_fac;
        p.event = EV_BOUNCE;
        return true;
    }
    return false;
}

}  // namespace

// Camera::new (camera.rs:22-72): aperture and focus lists fall back to [0]
// and [1] when either is empty or any focus distance is zero; each fov
// becomes its linear form tan(fov/2), computed as sin/cos.
Camera camera_new(const CameraDesc& d) {
    Camera c;
    c.transforms = d.transforms;
    c.apertures = d.aperture_radii;
    c.focuses = d.focus_distances;
    bool no_lens = c.apertures.empty() || c.focuses.empty();
    for (float f : c.focuses) no_lens = no_lens || f == 0.0f;
    if (no_lens) {
        c.apertures.assign(1, 0.0f);
        c.focuses.assign(1, 1.0f);
    }
    for (float f : d.fovs) c.tfovs.push_back(std::sin(f / 2.0f) / std::cos(f / 2.0f));
    return c;
}

// Camera::generate_ray (camera.rs:74-102): every list is lerped over the
// shutter on its own; (u, v) in [0, 1)^2 pick the point on the lens.
Ray generate_ray(const Camera& c, float x, float y, float time, float wavelength, float u, float v) {
    const Transform xf = lerp_slice(c.transforms.data(), c.transforms.size(), time);
    const float tfov = lerp_slice(c.tfovs.data(), c.tfovs.size(), time);
    const float aperture = lerp_slice(c.apertures.data(), c.apertures.size(), time);
    const float focus = lerp_slice(c.focuses.data(), c.focuses.size(), time);
    float lu, lv;
    square_to_circle((u * 2.0f) - 1.0f, (v * 2.0f) - 1.0f, lu, lv);
    const Vec3 orig(aperture * lu, aperture * lv, 0.0f);
    const Vec3 dir = normalized(Vec3((x * tfov) - (orig.x / focus), (y * tfov) - (orig.y / focus), 1.0f));
    return Ray{xform_point(orig, xf), xform_vector(dir, xf), time, wavelength, INFINITY};
}

// Renderer::render_job (renderer.rs:194-310): buckets of consecutive pixels
// in scanline order, at most 65,536 paths each, one tracer per thread.
void render(const RenderScene& sc, const RenderSettings& rs, float* xyz, uint32_t threads, double& trace_seconds,
            uint64_t& rays_out) {
    const uint32_t W = rs.width, H = rs.height, spp = rs.spp;
    const float cmpx = 1.0f / (float)W, cmpy = 1.0f / (float)H;
    const float min_x = -1.0f, max_x = 1.0f;
    const float min_y = -((float)H / (float)W), max_y = (float)H / (float)W;
    const float x_extent = max_x - min_x, y_extent = max_y - min_y;

    const Camera cam = camera_new(sc.camera);

    for (uint64_t i = 0; i < (uint64_t)W * H * 3; i++) xyz[i] = 0.0f;
    const uint64_t pixels = (uint64_t)W * H;
    const uint64_t bucket_px = std::max<uint64_t>(1, 65536 / std::max<uint32_t>(1, spp));
    std::atomic<uint64_t> next{0};
    std::vector<double> tsec(threads, 0.0);
    std::vector<uint64_t> nrays(threads, 0);

    auto worker = [&](uint32_t w) {
        Tracer tracer(sc.root);
        RayBatch rays;
        std::vector<LightPath> paths;
        for (;;) {
            uint64_t p0 = next.fetch_add(bucket_px);
            if (p0 >= pixels) break;
            uint64_t p1 = std::min(pixels, p0 + bucket_px);
            paths.clear();
            rays.clear();
            for (uint64_t pix = p0; pix < p1; pix++) {
                uint32_t x = (uint32_t)(pix % W), y = (uint32_t)(pix / W);
                for (uint32_t si = 0; si < spp; si++) {
                    float d0[4], d1[4];
                    get_sample_4d(si, 0, x, y, rs.seed, d0);
                    get_sample_4d(si, 1, x, y, rs.seed, d1);
                    float filter_x = probit(d0[3], 2.0f / 6.0f) + 0.5f;
                    float filter_y = probit(d1[0], 2.0f / 6.0f) + 0.5f;
                    float samp_x = (filter_x + (float)x) * cmpx;
                    float samp_y = (filter_y + (float)y) * cmpy;
                    float img_x = (samp_x - 0.5f) * x_extent, img_y = (0.5f - samp_y) * y_extent;
                    LightPath p;
                    p.sampling_seed = rs.seed ^ 0x40d4682bu;
                    p.px = x;
                    p.py = y;
                    p.sample_number = si;
                    p.time = d0[0];
                    p.wavelength = map_0_1_to_wavelength(golden_ratio_sample(si, hash_u32((x << 16) ^ y, rs.seed)));
                    const Ray ray = generate_ray(cam, img_x, img_y, p.time, p.wavelength, d0[1], d0[2]);
                    paths.push_back(p);
                    rays.push(ray, false);
                }
            }
            size_t pi = paths.size();
            while (pi > 0) {
                auto t0 = std::chrono::steady_clock::now();
                const std::vector<Isect>& is = tracer.trace(rays);
                tsec[w] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                size_t new_end = 0;
                for (size_t i = 0; i < pi; i++) {
                    if (path_next(paths[i], sc, is[i], rays, i, rs.max_bounces)) {
                        std::swap(paths[new_end], paths[i]);
                        std::swap(rays.hot[new_end], rays.hot[i]);
                        std::swap(rays.cold[new_end], rays.cold[i]);
                        new_end++;
                    }
                }
                rays.hot.resize(new_end);
                rays.cold.resize(new_end);
                pi = new_end;
            }
            // Film accumulation (renderer.rs:299-310): paths of a pixel in
            // sample order.
            std::sort(paths.begin(), paths.end(), [](const LightPath& a, const LightPath& b) {
                if (a.py != b.py) return a.py < b.py;
                if (a.px != b.px) return a.px < b.px;
                return a.sample_number < b.sample_number;
            });
            for (const LightPath& p : paths) {
                XYZ c = xyz_from_spectral_sample(p.color, p.wavelength);
                float* px = xyz + 3 * ((uint64_t)p.py * W + p.px);
                px[0] += c.x / (float)spp;
                px[1] += c.y / (float)spp;
                px[2] += c.z / (float)spp;
            }
        }
        nrays[w] = tracer.ray_trace_count;
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 0; w < threads; w++) pool.emplace_back(worker, w);
    for (auto& t : pool) t.join();
    trace_seconds = 0;
    rays_out = 0;
    for (uint32_t w = 0; w < threads; w++) {
        trace_seconds += tsec[w];
        rays_out += nrays[w];
    }
}

}  // namespace psyref
