// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Flat C interface over the restatement, loaded by the test suite (ctypes)
// and by the benchmark's CPU-baseline leg.  Never linked by the product.
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>

#include "psyref_scene.h"

using namespace psyref;

namespace {
struct OracleScene {
    Assembly root;
    bool built = false;
};
Transform xf_from(const float* v) { return Transform::from_values(v); }
void xf_to(const Transform& m, float* v) { m.to_values(v); }

// Canonical pre-order dump of a BVH4 (shared format with the product's dump):
//   internal: 0, child_count, traversal_code, bounds_len, bounds_len*24 f32 bits, children...
//   leaf:     1, obj_start, obj_end
void dump_node(const BVH4& bvh, size_t ni, std::vector<uint32_t>& out) {
    const BVH4Node& n = bvh.nodes[ni];
    if (n.leaf) {
        out.push_back(1);
        out.push_back((uint32_t)n.obj_start);
        out.push_back((uint32_t)n.obj_end);
        return;
    }
    out.push_back(0);
    out.push_back((uint32_t)n.children_len);
    out.push_back(n.traversal_code);
    out.push_back((uint32_t)n.bounds.size());
    for (const BBox4& b : n.bounds) {
        const float* arrs[6] = {b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax};
        for (int a = 0; a < 6; a++)
            for (int i = 0; i < 4; i++) out.push_back(f2u(arrs[a][i]));
    }
    for (size_t c = 0; c < n.children_len; c++) dump_node(bvh, n.children_start + c, out);
}
}  // namespace

extern "C" {

// ---- math hooks: the reference's unit tests are restated against these ----
float psyref_lerp_f32(float a, float b, float alpha) { return lerp(a, b, alpha); }
float psyref_lerp_slice_f32(const float* s, uint32_t n, float alpha) { return lerp_slice(s, n, alpha); }
void psyref_xform_lerp(const float* a, const float* b, float alpha, float* out) {
    xf_to(lerp(xf_from(a), xf_from(b), alpha), out);
}
void psyref_xform_mul(const float* a, const float* b, float* out) { xf_to(xform_compose(xf_from(a), xf_from(b)), out); }
void psyref_xform_inverse(const float* a, float* out) { xf_to(xform_inverse(xf_from(a)), out); }
void psyref_point_xform(const float* p, const float* m, float* out) {
    Vec3 r = xform_point(Vec3(p[0], p[1], p[2]), xf_from(m));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void psyref_vector_xform(const float* p, const float* m, float* out) {
    Vec3 r = xform_vector(Vec3(p[0], p[1], p[2]), xf_from(m));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void psyref_normal_xform(const float* p, const float* m, float* out) {
    Vec3 r = xform_normal(Vec3(p[0], p[1], p[2]), xf_from(m));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void psyref_vec_lerp(const float* a, const float* b, float alpha, float* out) {
    Vec3 r = lerp(Vec3(a[0], a[1], a[2]), Vec3(b[0], b[1], b[2]), alpha);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void psyref_vec_normalize(const float* a, float* out) {
    Vec3 r = normalized(Vec3(a[0], a[1], a[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float psyref_vec_length(const float* a) { return length(Vec3(a[0], a[1], a[2])); }
float psyref_vec_dot(const float* a, const float* b) { return dot(Vec3(a[0], a[1], a[2]), Vec3(b[0], b[1], b[2])); }
void psyref_vec_cross(const float* a, const float* b, float* out) {
    Vec3 r = cross(Vec3(a[0], a[1], a[2]), Vec3(b[0], b[1], b[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float psyref_increment_ulp(float v) { return increment_ulp(v); }
float psyref_decrement_ulp(float v) { return decrement_ulp(v); }
float psyref_fp_gamma(uint32_t n) { return fp_gamma(n); }
uint64_t psyref_log2_64(uint64_t n) { return log2_64(n); }
uint32_t psyref_hash_u32(uint32_t n, uint32_t seed) { return hash_u32(n, seed); }
uint64_t psyref_hash_u64(uint64_t n, uint64_t seed) { return hash_u64(n, seed); }
void psyref_quick_select_i32(int32_t* s, uint32_t n, uint32_t nth) { quick_select_i32(s, n, nth); }
void psyref_traversal_table(uint8_t* out) { std::memcpy(out, traversal_table(), 8 * 48); }

// Triangle test (triangle.rs:58).  out = (t, b0, b1, b2).  Returns 1 on hit.
int psyref_triangle_intersect(const float* orig, const float* dir, float max_t, const float* tri, float* out) {
    RayTriPrecompute pre(Vec3(dir[0], dir[1], dir[2]));
    TriHit h;
    bool hit = intersect_triangle(Vec3(orig[0], orig[1], orig[2]), pre, max_t, Vec3(tri[0], tri[1], tri[2]),
                                  Vec3(tri[3], tri[4], tri[5]), Vec3(tri[6], tri[7], tri[8]), h);
    if (hit) { out[0] = h.t; out[1] = h.b0; out[2] = h.b1; out[3] = h.b2; }
    return hit ? 1 : 0;
}

// BBox4 test (bbox4.rs:63).  b = xmin[4], xmax[4], ymin[4], ymax[4], zmin[4], zmax[4].
uint32_t psyref_bbox4_intersect(const float* b, const float* orig, const float* dir_inv, float max_t) {
    BBox4 bb;
    std::memcpy(bb.xmin, b + 0, 16); std::memcpy(bb.xmax, b + 4, 16);
    std::memcpy(bb.ymin, b + 8, 16); std::memcpy(bb.ymax, b + 12, 16);
    std::memcpy(bb.zmin, b + 16, 16); std::memcpy(bb.zmax, b + 20, 16);
    return bb.intersect_ray(Vec3(orig[0], orig[1], orig[2]), Vec3(dir_inv[0], dir_inv[1], dir_inv[2]), max_t);
}

// ---- scene -----------------------------------------------------------------
void* psyref_scene_new() { return new OracleScene(); }
void psyref_scene_free(void* s) { delete static_cast<OracleScene*>(s); }

// verts: [time_samples][vert_count][3]; normals: NULL or same layout.
int psyref_scene_add_mesh(void* s, const float* verts, uint32_t vert_count, uint32_t time_samples,
                          const float* normals, const uint32_t* tris, uint32_t tri_count) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    sc->root.objects.push_back(
        TriangleMesh::from_verts_and_indices(verts, vert_count, time_samples, normals, tris, tri_count));
    return (int)sc->root.objects.size() - 1;
}

// xforms: [n][12] values a..l of Transform::new_from_values (world -> object).
int psyref_scene_add_instance(void* s, uint32_t mesh, int32_t shader, const float* xforms, uint32_t n_xf) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    std::vector<Transform> xf;
    for (uint32_t i = 0; i < n_xf; i++) xf.push_back(xf_from(xforms + 12 * i));
    sc->root.add_instance(false, mesh, shader, xf.data(), n_xf);
    return (int)sc->root.instances.size() - 1;
}

// Nested assemblies (parse/psy_assembly.rs:26-37, AssemblyBuilder::
// add_assembly, scene/assembly.rs:204-217): moves the built root assembly of
// `child` into `parent` as a sub-assembly (and frees `child`); returns its
// index for psyref_scene_add_assembly_instance.
int psyref_scene_add_assembly(void* parent, void* child) {
    OracleScene* p = static_cast<OracleScene*>(parent);
    OracleScene* c = static_cast<OracleScene*>(child);
    if (!c->built) return -1;
    auto a = std::make_unique<Assembly>(std::move(c->root));
    p->root.assemblies.push_back(std::move(a));
    delete c;
    return (int)p->root.assemblies.size() - 1;
}

int psyref_scene_add_assembly_instance(void* s, uint32_t assembly, const float* xforms, uint32_t n_xf) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    if (assembly >= sc->root.assemblies.size()) return -1;
    std::vector<Transform> xf;
    for (uint32_t i = 0; i < n_xf; i++) xf.push_back(xf_from(xforms + 12 * i));
    sc->root.add_instance(true, assembly, -1, xf.data(), n_xf);
    return (int)sc->root.instances.size() - 1;
}

int psyref_scene_build(void* s) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    sc->root.build();
    sc->built = true;
    return 0;
}

// Canonical dump of a mesh's BVH4 (or the instance BVH when mesh < 0) and of
// the object order in its leaves.  Returns the number of u32 written, or the
// required size when `cap` is too small.
uint64_t psyref_scene_dump_bvh(void* s, int32_t mesh, uint32_t* out, uint64_t cap, uint32_t* order, uint64_t ocap) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    const BVH4& bvh = mesh < 0 ? sc->root.object_accel : sc->root.objects[mesh]->accel;
    std::vector<uint32_t> v;
    v.push_back((uint32_t)bvh.has_root);
    v.push_back((uint32_t)bvh.depth);
    v.push_back((uint32_t)bvh.node_count);
    if (bvh.has_root) dump_node(bvh, 0, v);
    if (v.size() <= cap) std::memcpy(out, v.data(), v.size() * 4);
    if (order) {
        if (mesh < 0) {
            for (size_t i = 0; i < sc->root.instances.size() && i < ocap; i++)
                order[i] = (uint32_t)sc->root.instances[i].id;
        } else {
            const auto& idx = sc->root.objects[mesh]->indices;
            for (size_t i = 0; i < idx.size() && i < ocap; i++) order[i] = idx[i].orig;
        }
    }
    return v.size();
}

// Traces `n` rays (SoA) in batches of `batch` rays (<= 65536) on `threads`
// worker threads, one Tracer per thread, as Renderer::render_job does
// (renderer.rs:194-310).  Only the trace calls are timed; `seconds` receives
// the summed per-thread trace time divided by the thread count.
int psyref_trace(void* s, const float* ox, const float* oy, const float* oz, const float* dx, const float* dy,
                 const float* dz, const float* time, const float* max_t, const uint8_t* occl, uint64_t n,
                 uint32_t batch, uint32_t threads, uint8_t* kind, float* t, float* b0, float* b1, float* b2,
                 uint32_t* prim_id, uint32_t* inst_id, float* nor_g /*n*3 or NULL*/, float* pos /*n*4 or NULL*/,
                 float* geom /*n*24 or NULL*/, uint64_t* node_tests_out, double* seconds) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    if (!sc->built || batch == 0 || batch > 65536) return -1;
    if (threads == 0) threads = 1;
    std::atomic<uint64_t> next{0};
    std::atomic<uint64_t> total_tests{0};
    std::vector<double> tsec(threads, 0.0);
    auto worker = [&](uint32_t w) {
        Tracer tracer(&sc->root);
        RayBatch rays;
        for (;;) {
            uint64_t start = next.fetch_add(batch);
            if (start >= n) break;
            uint64_t end = std::min<uint64_t>(n, start + batch);
            rays.clear();
            for (uint64_t i = start; i < end; i++) {
                Ray r;
                r.orig = Vec3(ox[i], oy[i], oz[i]);
                r.dir = Vec3(dx[i], dy[i], dz[i]);
                r.time = time[i];
                r.wavelength = 0.0f;
                r.max_t = max_t[i];
                rays.push(r, occl && occl[i] != 0);
            }
            auto t0 = std::chrono::steady_clock::now();
            const std::vector<Isect>& is = tracer.trace(rays);
            auto t1 = std::chrono::steady_clock::now();
            tsec[w] += std::chrono::duration<double>(t1 - t0).count();
            for (uint64_t i = start; i < end; i++) {
                const Isect& h = is[i - start];
                kind[i] = h.kind;
                t[i] = h.t; b0[i] = h.b0; b1[i] = h.b1; b2[i] = h.b2;
                prim_id[i] = h.prim_id;
                inst_id[i] = h.inst_id;
                if (nor_g) { nor_g[3 * i] = h.nor_g.x; nor_g[3 * i + 1] = h.nor_g.y; nor_g[3 * i + 2] = h.nor_g.z; }
                if (pos) { pos[4 * i] = h.pos.x; pos[4 * i + 1] = h.pos.y; pos[4 * i + 2] = h.pos.z; pos[4 * i + 3] = h.pos_err; }
                if (geom) {
                    // SurfaceIntersectionData: pos, pos_err, nor, nor_g,
                    // local_space (a..l), sample_pdf, 0 (psy_hit_geometry).
                    float* g = geom + 24 * i;
                    for (int k = 0; k < 24; k++) g[k] = 0.0f;
                    if (h.kind == ISECT_HIT) {
                        g[0] = h.pos.x; g[1] = h.pos.y; g[2] = h.pos.z; g[3] = h.pos_err;
                        g[4] = h.nor.x; g[5] = h.nor.y; g[6] = h.nor.z;
                        g[7] = h.nor_g.x; g[8] = h.nor_g.y; g[9] = h.nor_g.z;
                        h.local_space.to_values(g + 10);
                        g[22] = h.sample_pdf;
                    }
                }
            }
        }
        total_tests += tracer.node_tests;
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 0; w < threads; w++) pool.emplace_back(worker, w);
    for (auto& th : pool) th.join();
    double sum = 0;
    for (double v : tsec) sum += v;
    if (seconds) *seconds = sum / threads;
    if (node_tests_out) *node_tests_out = total_tests.load();
    return 0;
}

}  // extern "C"

// ---- rendering ---------------------------------------------------------------
#include "psyref_render.h"

namespace {
struct PsyrefColor { uint32_t kind; float a, b, c; };
struct PsyrefShader { uint32_t kind; PsyrefColor color; float roughness, fresnel, ior; };
struct PsyrefDistant { float radius; float dir[3]; PsyrefColor color; };
// Time samples of a distant light (counts 0: the PsyrefDistant values).
struct PsyrefDistantSamples {
    uint32_t n_radius; const float* radii;
    uint32_t n_direction; const float* directions;  // xyz per sample
    uint32_t n_color; const PsyrefColor* colors;
};
Color to_color(const PsyrefColor& c) {
    Color r;
    r.kind = c.kind;
    r.a = c.a; r.b = c.b; r.c = c.c;
    return r;
}
}  // namespace

extern "C" {

int psyref_scene_add_rect_light(void* s, float w, float h, const PsyrefColor* color) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    RectangleLight l;
    l.w = w; l.h = h; l.color = to_color(*color);
    sc->root.rect_lights.push_back(l);
    return (int)sc->root.rect_lights.size() - 1;
}
int psyref_scene_add_sphere_light(void* s, float radius, const PsyrefColor* color) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    SphereLight l;
    l.radius = radius; l.color = to_color(*color);
    sc->root.sphere_lights.push_back(l);
    return (int)sc->root.sphere_lights.size() - 1;
}
// lerp_slice over a colour list (the lights' colour at a time).
void psyref_color_lerp_slice(const PsyrefColor* cs, uint32_t n, float time, PsyrefColor* out) {
    std::vector<Color> v;
    for (uint32_t i = 0; i < n; i++) v.push_back(to_color(cs[i]));
    const Color c = lerp_slice(v.data(), v.size(), time);
    out->kind = c.kind;
    out->a = c.a;
    out->b = c.b;
    out->c = c.c;
}

// Time samples of light `light` of `kind` (1 rectangle: n_a (w, h) pairs;
// 2 sphere: n_a radii) and its colours; a zero count keeps the single
// value.  Call before psyref_scene_build.
int psyref_scene_set_light_samples(void* s, uint32_t kind, uint32_t light, uint32_t n_a, const float* a,
                                   uint32_t n_color, const PsyrefColor* colors) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    std::vector<Color> cs;
    for (uint32_t i = 0; i < n_color; i++) cs.push_back(to_color(colors[i]));
    if (kind == OBJ_RECT_LIGHT && light < sc->root.rect_lights.size()) {
        sc->root.rect_lights[light].dims_t.assign(a, a + 2 * n_a);
        sc->root.rect_lights[light].colors_t = cs;
        return 0;
    }
    if (kind == OBJ_SPHERE_LIGHT && light < sc->root.sphere_lights.size()) {
        sc->root.sphere_lights[light].radii_t.assign(a, a + n_a);
        sc->root.sphere_lights[light].colors_t = cs;
        return 0;
    }
    return -1;
}
// kind: 1 rectangle light, 2 sphere light.  Lights need transforms
// (the reference lerps the transform list, rectangle_light.rs:270).
int psyref_scene_add_light_instance(void* s, uint32_t kind, uint32_t light, const float* xforms, uint32_t n_xf) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    if (n_xf == 0 || (kind != OBJ_RECT_LIGHT && kind != OBJ_SPHERE_LIGHT)) return -1;
    std::vector<Transform> xf;
    for (uint32_t i = 0; i < n_xf; i++) xf.push_back(xf_from(xforms + 12 * i));
    sc->root.add_instance(false, light, -1, xf.data(), n_xf, (uint8_t)kind);
    return (int)sc->root.instances.size() - 1;
}

// Renders the built scene.  camera_xforms: n_cam x 12 (camera to world);
// fovs (radians), apertures and focus distances: time samples of the lens,
// as parse_camera collects them (an empty aperture or focus list disables
// the lens).
int psyref_render(void* s, const PsyrefShader* shaders, uint32_t n_shaders, const PsyrefColor* background,
                  const PsyrefDistant* distant, uint32_t n_distant, const PsyrefDistantSamples* distant_samples,
                  const float* camera_xforms, uint32_t n_cam,
                  const float* fovs, uint32_t n_fov, const float* apertures, uint32_t n_ap, const float* focus,
                  uint32_t n_fd, uint32_t width, uint32_t height, uint32_t spp,
                  uint32_t max_bounces, uint32_t seed, uint32_t threads, float* xyz, double* trace_seconds,
                  uint64_t* rays) {
    OracleScene* sc = static_cast<OracleScene*>(s);
    if (!sc->built || n_cam == 0 || n_fov == 0 || width == 0 || height == 0 || spp == 0) return -1;
    RenderScene rsc;
    rsc.root = &sc->root;
    for (uint32_t i = 0; i < n_shaders; i++) {
        Closure c;
        c.kind = shaders[i].kind;
        c.color = to_color(shaders[i].color);
        c.roughness = shaders[i].roughness;
        c.fresnel = shaders[i].fresnel;
        c.ior = shaders[i].ior;
        rsc.shaders.push_back(c);
    }
    for (const Instance& inst : sc->root.instances)
        if (inst.obj_kind == OBJ_MESH && inst.shader >= 0 && (uint32_t)inst.shader >= n_shaders) return -2;
    rsc.world.background = to_color(*background);
    for (uint32_t i = 0; i < n_distant; i++) {
        DistantDiskLight l;
        l.radius = distant[i].radius;
        l.direction = Vec3(distant[i].dir[0], distant[i].dir[1], distant[i].dir[2]);
        l.color = to_color(distant[i].color);
        if (distant_samples) {
            const PsyrefDistantSamples& t = distant_samples[i];
            l.radii_t.assign(t.radii, t.radii + t.n_radius);
            for (uint32_t k = 0; k < t.n_direction; k++)
                l.directions_t.push_back(Vec3(t.directions[3 * k], t.directions[3 * k + 1], t.directions[3 * k + 2]));
            for (uint32_t k = 0; k < t.n_color; k++) l.colors_t.push_back(to_color(t.colors[k]));
        }
        rsc.world.lights.push_back(l);
    }
    for (uint32_t i = 0; i < n_cam; i++) rsc.camera.transforms.push_back(xf_from(camera_xforms + 12 * i));
    rsc.camera.fovs.assign(fovs, fovs + n_fov);
    rsc.camera.aperture_radii.assign(apertures, apertures + n_ap);
    rsc.camera.focus_distances.assign(focus, focus + n_fd);
    rsc.build_lights();
    RenderSettings rs;
    rs.width = width; rs.height = height; rs.spp = spp; rs.max_bounces = max_bounces; rs.seed = seed;
    double ts = 0;
    uint64_t nr = 0;
    render(rsc, rs, xyz, threads ? threads : 1, ts, nr);
    if (trace_seconds) *trace_seconds = ts;
    if (rays) *rays = nr;
    return 0;
}

// Camera::new then Camera::generate_ray for image-plane point (x, y):
// out = origin xyz, direction xyz (world space).
int psyref_camera_ray(const float* xforms, uint32_t n_xf, const float* fovs, uint32_t n_fov, const float* apertures,
                      uint32_t n_ap, const float* focus, uint32_t n_fd, float x, float y, float time, float u, float v,
                      float* out) {
    if (n_xf == 0 || n_fov == 0) return -1;
    CameraDesc d;
    for (uint32_t i = 0; i < n_xf; i++) d.transforms.push_back(xf_from(xforms + 12 * i));
    d.fovs.assign(fovs, fovs + n_fov);
    d.aperture_radii.assign(apertures, apertures + n_ap);
    d.focus_distances.assign(focus, focus + n_fd);
    const Ray r = generate_ray(camera_new(d), x, y, time, 550.0f, u, v);
    out[0] = r.orig.x; out[1] = r.orig.y; out[2] = r.orig.z;
    out[3] = r.dir.x; out[4] = r.dir.y; out[5] = r.dir.z;
    return 0;
}

float psyref_halton(uint32_t dim, uint32_t index) { return halton().sample(dim, index); }
void psyref_sample_4d(uint32_t i, uint32_t set, uint32_t px, uint32_t py, uint32_t seed, float* out) {
    get_sample_4d(i, set, px, py, seed, out);
}
float psyref_probit(float n, float width) { return probit(n, width); }
float psyref_fast_exp(float p) { return fast_exp(p); }
void psyref_xyz_to_spectrum(float x, float y, float z, float hero, float* out) {
    Vec4 r = xyz_to_spectrum_4(x, y, z, wavelengths(hero));
    for (int k = 0; k < 4; k++) out[k] = r[k];
}
void psyref_spectrum_to_xyz(const float* e, float hero, float* out) {
    XYZ r = xyz_from_spectral_sample(Vec4(e[0], e[1], e[2], e[3]), hero);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float psyref_planck(float temperature, float wavelength) { return plancks_law(temperature, wavelength); }
const char* psyref_meng_sha256() { return PSY_MENG_TABLES_SHA256; }

}  // extern "C"

// ---- single-function hooks that the mathematical checks of the closures,
// light samplers and light tree run against --------------------------------
namespace {
Closure to_closure(const PsyrefShader* sh) {
    Closure c;
    c.kind = sh->kind;
    c.color = to_color(sh->color);
    c.roughness = sh->roughness;
    c.fresnel = sh->fresnel;
    c.ior = sh->ior;
    return c;
}
Vec3 v3at(const float* p, uint32_t i) { return Vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
void put3(float* p, uint32_t i, Vec3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
void put4(float* p, uint32_t i, Vec4 v) { for (int k = 0; k < 4; k++) p[4 * i + k] = v[k]; }
}  // namespace

extern "C" {

// SurfaceClosure::evaluate (surface_closure.rs:57-155) for n (inc, out, nor,
// nor_g) quadruples, 3 floats each; filter is 4 floats per item.
void psyref_closure_evaluate_n(const PsyrefShader* sh, uint32_t n, const float* inc, const float* out,
                               const float* nor, const float* nor_g, float wl, float* filter, float* pdf) {
    const Closure c = to_closure(sh);
    for (uint32_t i = 0; i < n; i++) {
        Vec4 f;
        float p;
        closure_evaluate(c, v3at(inc, i), v3at(out, i), v3at(nor, i), v3at(nor_g, i), wl, f, p);
        put4(filter, i, f);
        pdf[i] = p;
    }
}

// SurfaceClosure::sample for n (inc, nor, nor_g, (u, v)) items; `w` (NULL:
// zeros) is the dielectric's third number.
void psyref_closure_sample_n(const PsyrefShader* sh, uint32_t n, const float* inc, const float* nor,
                             const float* nor_g, const float* uv, const float* w, float wl, float* out, float* filter,
                             float* pdf) {
    const Closure c = to_closure(sh);
    for (uint32_t i = 0; i < n; i++) {
        Vec3 o;
        Vec4 f;
        float p;
        closure_sample(c, v3at(inc, i), v3at(nor, i), v3at(nor_g, i), uv[2 * i], uv[2 * i + 1], w ? w[i] : 0.0f, wl,
                       o, f, p);
        put3(out, i, o);
        put4(filter, i, f);
        pdf[i] = p;
    }
}

// SurfaceClosure::estimate_eval_over_sphere_light for n items.
void psyref_closure_estimate_n(const PsyrefShader* sh, uint32_t n, const float* inc, const float* to_light,
                               const float* r2, const float* nor, const float* nor_g, float* out) {
    const Closure c = to_closure(sh);
    for (uint32_t i = 0; i < n; i++)
        out[i] = closure_estimate_sphere_light(c, v3at(inc, i), v3at(to_light, i), r2[i], v3at(nor, i),
                                               v3at(nor_g, i));
}

// SphereLight::sample_from_point (sphere_light.rs:85-192) with the light's
// world-to-object transform `xform` (12 floats), for n points and (u, v).
void psyref_sphere_light_sample_n(float radius, const PsyrefColor* color, const float* xform, uint32_t n,
                                  const float* arr, const float* uv, float wl, float* point, float* normal,
                                  float* pdf, float* err, float* col) {
    SphereLight l;
    l.radius = radius;
    l.color = to_color(*color);
    const Transform x = xf_from(xform);
    for (uint32_t i = 0; i < n; i++) {
        const LightSample s = l.sample_from_point(x, v3at(arr, i), uv[2 * i], uv[2 * i + 1], wl);
        put3(point, i, s.point);
        put3(normal, i, s.normal);
        pdf[i] = s.pdf;
        err[i] = s.point_err;
        put4(col, i, s.color);
    }
}

// SphereLight::sample_pdf (sphere_light.rs:51-81) for n points.
void psyref_sphere_light_pdf_n(float radius, const float* xform, uint32_t n, const float* arr, float* pdf) {
    SphereLight l;
    l.radius = radius;
    const Transform x = xf_from(xform);
    for (uint32_t i = 0; i < n; i++) pdf[i] = l.sample_pdf(x, v3at(arr, i));
}

// DistantDiskLight::sample_from_point (distant_disk_light.rs:59-85) for n
// (u, v) pairs: colour (4 floats), shadow direction, pdf.
void psyref_distant_light_sample_n(float radius, const float* dir, const PsyrefColor* color, uint32_t n,
                                   const float* uv, float wl, float* col, float* shadow, float* pdf) {
    DistantDiskLight l;
    l.radius = radius;
    l.direction = Vec3(dir[0], dir[1], dir[2]);
    l.color = to_color(*color);
    for (uint32_t i = 0; i < n; i++) {
        Vec4 c;
        Vec3 sv;
        float p;
        l.sample_from_point(uv[2 * i], uv[2 * i + 1], wl, c, sv, p);
        put4(col, i, c);
        put3(shadow, i, sv);
        pdf[i] = p;
    }
}

// A light tree (light_tree.rs:211-372) over n lights with one box each
// (6 floats: min xyz, max xyz) and an energy each, then LightTree::select
// (light_tree.rs:135-201) at one shading point for m numbers `us`: the
// chosen light (its index in the input), the selection probability and the
// whittled number.  Returns 0, or -1 when n is 0.
int psyref_light_tree_select_n(uint32_t n, const float* boxes, const float* energies, const PsyrefShader* sh,
                               const float* inc, const float* pos, const float* nor, const float* nor_g, float time,
                               uint32_t m, const float* us, uint32_t* index, float* prob, float* whittled) {
    if (n == 0) return -1;
    std::vector<BBox> bb(n);
    for (uint32_t i = 0; i < n; i++)
        bb[i] = BBox(Vec3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]),
                     Vec3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]));
    std::vector<uint32_t> keys(n);
    for (uint32_t i = 0; i < n; i++) keys[i] = i;
    LightTree t;
    t.build(keys.data(), n, [&](uint32_t k) { return BoundsSlice{&bb[k], 1}; },
            [&](uint32_t k) { return energies[k]; });
    const Closure c = to_closure(sh);
    for (uint32_t i = 0; i < m; i++) {
        size_t idx = 0;
        float p = 0, w = 0;
        t.select(v3at(inc, 0), v3at(pos, 0), v3at(nor, 0), v3at(nor_g, 0), c, time, us[i], idx, p, w);
        index[i] = keys[idx];
        prob[i] = p;
        whittled[i] = w;
    }
    return 0;
}

// Pre-order dump of the same light tree: per node (leaf, light or child
// count, energy bits, box: 6 floats as bits), 9 u32 each.  Returns the node
// count; writes at most `cap` nodes.
uint32_t psyref_light_tree_dump(uint32_t n, const float* boxes, const float* energies, uint32_t* out, uint32_t cap) {
    if (n == 0) return 0;
    std::vector<BBox> bb(n);
    for (uint32_t i = 0; i < n; i++)
        bb[i] = BBox(Vec3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]),
                     Vec3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]));
    std::vector<uint32_t> keys(n);
    for (uint32_t i = 0; i < n; i++) keys[i] = i;
    LightTree t;
    t.build(keys.data(), n, [&](uint32_t k) { return BoundsSlice{&bb[k], 1}; },
            [&](uint32_t k) { return energies[k]; });
    uint32_t count = 0;
    std::function<void(const LightTree::Node&)> walk = [&](const LightTree::Node& nd) {
        if (count < cap) {
            uint32_t* o = out + 9 * count;
            o[0] = nd.leaf ? 1u : 0u;
            o[1] = nd.leaf ? keys[nd.light_index] : (uint32_t)nd.children.size();
            o[2] = f2u(nd.energy);
            const BBox b = nd.bounds.front();
            o[3] = f2u(b.min.x); o[4] = f2u(b.min.y); o[5] = f2u(b.min.z);
            o[6] = f2u(b.max.x); o[7] = f2u(b.max.y); o[8] = f2u(b.max.z);
        }
        count++;
        for (const LightTree::Node& ch : nd.children) walk(ch);
    };
    walk(t.root);
    return count;
}

}  // extern "C"
