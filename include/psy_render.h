/*
 * psy_render.h — C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * Replaces the renderer's per-thread bucket loop above the tracer:
 *   Renderer::render(..) -> (Image, RenderStats)      (src/renderer.rs:76-192)
 *   Renderer::render_job / LightPath::{new,next}      (src/renderer.rs:194-692)
 *   Camera::generate_ray                              (src/camera.rs:74-102)
 *   Scene::sample_lights / Assembly::sample_lights    (src/scene/mod.rs:28-101,
 *                                                      src/scene/assembly.rs:46-141)
 * with the scene's shaders (SimpleSurfaceShader, src/shading/mod.rs:28-61)
 * and world (World, src/scene/world.rs:4-8) given at renderer creation and
 * the camera (Camera::new, src/camera.rs:22-72) and render settings
 * (Renderer, src/renderer.rs:32-38) per render.  The output is the XYZ image
 * the reference accumulates before writing a file (Image, src/image.rs).
 *
 * Light paths follow the reference's state machine exactly (hero
 * wavelengths, next-event estimation through the light tree, power-heuristic
 * MIS), with the Faure-permuted Halton sequence in place of the Sobol
 * sampler for the per-pixel sample dimensions.  Paths are processed as
 * device-resident wavefronts: camera, closest-hit and shadow traversal,
 * material-sorted shading, and an in-order film accumulation.
 */
#ifndef PSY_RENDER_H
#define PSY_RENDER_H

#include "psy_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psy_renderer psy_renderer;

/* Surface shaders (SimpleSurfaceShader, src/shading/mod.rs:28-40).
 * PSY_SHADER_DIELECTRIC has no counterpart in the reference, whose closures
 * are Lambert, GGX reflection and emission (surface_closure.rs:19-30): a
 * rough dielectric interface, GGX microfacet reflection and refraction with
 * exact Fresnel reflectance (Walter et al., "Microfacet Models for
 * Refraction through Rough Surfaces", EGSR 2007), in throughput form (no
 * (eta_i/eta_o)^2 radiance factor).  `roughness` is its GGX alpha in
 * [1e-4, 1], `ior` in (1.0001, 10] the index of refraction on the side the
 * geometric normal points away from, and `color` tints refracted light. */
#define PSY_SHADER_LAMBERT 0u
#define PSY_SHADER_GGX 1u
#define PSY_SHADER_EMIT 2u
#define PSY_SHADER_DIELECTRIC 3u
typedef struct {
    uint32_t kind;
    psy_color color;
    float roughness;  /* GGX and dielectric */
    float fresnel;    /* GGX only */
    float ior;        /* dielectric only */
} psy_shader_desc;

/* Distant disk light (DistantDiskLight::new,
 * src/light/distant_disk_light.rs:24-36): angular radius in radians, the
 * direction the light travels, colour.  Each may be time samples, as for
 * the lights of psy_rt.h: a non-zero count replaces the value before it. */
typedef struct {
    float radius;
    float direction[3];
    psy_color color;
    uint32_t radius_count;       /* 0 = radius above */
    const float* radii;
    uint32_t direction_count;    /* 0 = direction above */
    const float* directions;     /* xyz per sample */
    uint32_t color_count;        /* 0 = color above */
    const psy_color* colors;
} psy_distant_light_desc;

/* Shaders (indexed by psy_instance_desc.shader) and world. */
typedef struct {
    uint32_t shader_count;
    const psy_shader_desc* shaders;
    psy_color background;
    uint32_t distant_light_count;
    const psy_distant_light_desc* distant_lights;
} psy_world_desc;

/* Camera (Camera::new, src/camera.rs:22-72): camera-to-world transforms
 * (12 floats each, as in psy_scene_desc), horizontal field of view in
 * radians, thin-lens aperture radius and focus distance (aperture 0 gives a
 * pinhole; focus distance 0 disables the lens as in the reference).
 *
 * The lens parameters may be time samples too (Camera::generate_ray,
 * camera.rs:74-79): a non-zero count replaces the scalar before it by that
 * many samples, interpolated over the shutter like the transforms
 * (lerp_slice, src/lerp.rs:20-35), each list on its own; the field of view
 * is interpolated in its linear form tan(fov/2).  A zero focus distance
 * among the samples disables the lens for the whole render.  At most 64
 * samples per list. */
typedef struct {
    uint32_t xform_count;
    const float* xforms;
    float fov;
    float aperture_radius;
    float focus_distance;
    uint32_t fov_count;             /* 0 = the scalar fov */
    const float* fovs;
    uint32_t aperture_count;        /* 0 = the scalar aperture_radius */
    const float* aperture_radii;
    uint32_t focus_count;           /* 0 = the scalar focus_distance */
    const float* focus_distances;
} psy_camera_desc;

/* One render (Renderer, src/renderer.rs:32-38): resolution, samples per
 * pixel, seed, and the bounce limit of a path (the reference fixes it at 2,
 * renderer.rs:599).  A crop region (crop_width > 0) renders only the pixels
 * [crop_x, crop_x + crop_width) x [crop_y, crop_y + crop_height) of the
 * width x height image, which keeps its size with the rest black, as the
 * reference's --crop does (renderer.rs:97-108); pixels inside are the same
 * as in the full render. */
typedef struct {
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t max_bounces;
    uint32_t seed;
    uint32_t crop_x, crop_y, crop_width, crop_height;
} psy_render_settings;

/* RenderStats (src/renderer.rs:41-49), as measured on the device. */
typedef struct {
    uint64_t paths;        /* camera paths (width * height * spp) */
    uint64_t rays;         /* rays traced: closest-hit and shadow */
    uint64_t waves;        /* path wavefronts the render was split into */
    double seconds;        /* wall time of the render call */
    double trace_ms;       /* traversal kernel time (HIP events) */
} psy_render_stats;

/* Creates a renderer for `scene` (which must outlive it): uploads the scene
 * if needed, the shaders, the world and the light tree, and allocates path
 * state for up to `max_paths` concurrent paths (0 = a default of 4M).
 * Every mesh instance must name a shader in range. */
int psy_renderer_create(const psy_scene* scene, const psy_world_desc* world, uint64_t max_paths,
                        psy_renderer** out);
int psy_renderer_destroy(psy_renderer* renderer);

/* Renders the image: width*height XYZ triples, row 0 at the top, each pixel
 * the mean over its samples (accumulated in sample order).  `mem` says
 * whether `xyz` is a host (PSY_MEM_HOST) or device (PSY_MEM_DEVICE)
 * pointer.  Synchronous.  `stats` may be NULL. */
int psy_render(psy_renderer* renderer, const psy_camera_desc* camera, const psy_render_settings* settings,
               float* xyz, uint32_t mem, psy_render_stats* stats);

/* Launch shape of the renderer's traversal kernel (see psy_trace_tuning;
 * the renderer always uses the persistent kernel). */
int psy_renderer_get_tuning(const psy_renderer* renderer, psy_trace_tuning* tuning);
int psy_renderer_set_tuning(psy_renderer* renderer, const psy_trace_tuning* tuning);

/* Execution options, none of which changes results: PSY_RENDER_NO_GRAPHS
 * launches every wave's kernels directly instead of replaying a captured
 * hipGraph; PSY_RENDER_NO_SORT shades closest hits in queue order instead of
 * binning them by closure kind first; PSY_RENDER_PLAIN renders in one
 * kernel launch over the whole frame (one lane per pixel tracing and
 * shading its samples' paths to the end, no wavefronts), the baseline the
 * wavefront and served forms are measured against.
 * PSY_RENDER_SHADE_WAVES(n), n = 1..4, selects the build of the shading
 * kernel compiled for at least n waves per SIMD (a register budget of 512/n
 * per lane; tighter budgets spill to scratch): of the one-launch kernel
 * when given with PSY_RENDER_PLAIN, of the wavefront's shading kernel
 * otherwise; 0 in those bits keeps the current choice (2 for both). */
#define PSY_RENDER_NO_GRAPHS 1u
#define PSY_RENDER_NO_SORT 2u
#define PSY_RENDER_PLAIN 4u
/* Measurement options of the wavefront form (no counterpart in the
 * reference): PSY_RENDER_PROFILE launches every wave directly with HIP
 * events around each kernel and tallies per-stage device time and ray
 * counts (psy_renderer_get_profile); PSY_RENDER_COUNT traces with the
 * instrumented traversal kernel, tallying node visits and triangle tests
 * (slower; images are unchanged). */
#define PSY_RENDER_PROFILE 16u
#define PSY_RENDER_COUNT 32u
#define PSY_RENDER_SHADE_WAVES_SHIFT 8u
#define PSY_RENDER_SHADE_WAVES_MASK (7u << PSY_RENDER_SHADE_WAVES_SHIFT)
#define PSY_RENDER_SHADE_WAVES(n) ((uint32_t)(n) << PSY_RENDER_SHADE_WAVES_SHIFT)
int psy_renderer_set_options(psy_renderer* renderer, uint32_t flags);

/* Tallies of the wavefront renders since the last call, with
 * PSY_RENDER_PROFILE (stage times, paths, rays) and PSY_RENDER_COUNT
 * (traversal counters of closest-hit and shadow rays); the call resets
 * them.  Stage times are device milliseconds between HIP events around each
 * kernel: camera_kernel, the closest-hit and shadow traversals, shadow
 * results, the material sort, shade_kernel, and the film kernels. */
typedef struct {
    double camera_ms, trace_ms, shadow_trace_ms, shadow_ms, sort_ms, shade_ms, film_ms;
    uint64_t paths, closest_rays, shadow_rays;
    uint64_t closest_node_visits, closest_tri_tests, shadow_node_visits, shadow_tri_tests;
    uint64_t motion_node_visits, motion_tri_tests, instance_visits;
} psy_render_profile;
int psy_renderer_get_profile(psy_renderer* renderer, psy_render_profile* profile);

/* Autotuner: renders `trial` (a reduced version of the intended renders)
 * with every candidate launch shape of the traversal kernel (waves per CU,
 * refill threshold, LDS stack depth, register budget), then its descent
 * threshold (leaf_min), then every shading register budget,
 * then with and without material sorting, then the one-launch form with each
 * of its register budgets, keeps the fastest on the renderer and returns it
 * (either output may be NULL; `best_flags` holds PSY_RENDER_NO_SORT if
 * sorting lost, PSY_RENDER_PLAIN if the one-launch form won, and
 * PSY_RENDER_SHADE_WAVES(n) of the chosen form's budget).  The renderer keeps
 * the best budget of each form, so a later set_options switching form without
 * a budget uses it.  Images do not depend on the choice; the request server
 * always batches in waves. */
int psy_renderer_autotune(psy_renderer* renderer, const psy_camera_desc* camera, const psy_render_settings* trial,
                          psy_trace_tuning* best_tuning, uint32_t* best_flags);

/* ---- request server ---------------------------------------------------
 * Serves render requests against several renderers (one device-resident
 * scene each) with continuous batching.  Requests queue on submission and
 * are admitted in arrival order once the HBM budget has room for their film
 * and their renderer a free request slot; admitted requests keep path state
 * and film on the device.  A worker thread fills each renderer's waves with
 * the next paths of its active requests, so requests that arrive while
 * others render join the next wave, and the renderers' streams overlap on
 * the device.  Each image equals what psy_render gives for the request
 * alone.  Renderers given to a server must not be used otherwise until the
 * server is destroyed.  There is no counterpart in the reference, which
 * renders one scene per process (src/main.rs). */
typedef struct psy_server psy_server;

typedef struct {
    uint64_t hbm_budget;   /* bytes for scenes, path state and films; 0 = no limit */
    uint32_t max_active;   /* admitted requests per renderer, 1..64; 0 = 64 */
} psy_server_config;

typedef struct {
    uint64_t paths;        /* width * height * spp */
    uint64_t waves;        /* waves that carried paths of the request */
    double queued_ms;      /* submission to admission */
    double latency_ms;     /* submission to image complete on the device */
} psy_request_stats;

typedef struct {
    uint64_t waves;        /* waves launched */
    uint64_t paths;        /* paths rendered */
    uint64_t queued;       /* requests waiting for admission */
    uint64_t fixed_bytes;  /* scenes and path state */
    uint64_t film_bytes;   /* films held (admitted, uncollected or pooled) */
} psy_server_stats;

int psy_server_create(psy_renderer* const* renderers, uint32_t count, const psy_server_config* config,
                      psy_server** out);
/* Releases the server: requests not yet admitted are dropped, no further
 * waves start, and the waves already on the device finish first.  Results
 * not collected by psy_server_wait are discarded. */
int psy_server_destroy(psy_server* server);
/* Queues a render of `renderer` (an index into the creation list). */
int psy_server_submit(psy_server* server, uint32_t renderer, const psy_camera_desc* camera,
                      const psy_render_settings* settings, uint64_t* request_id);
/* Blocks until the request is complete, copies its XYZ image to host memory
 * (`xyz` may be NULL to discard it) and releases the request.  A request
 * whose film cannot be allocated while no other request runs (so no memory
 * will free up) completes with PSY_E_NOMEM.  A psy_server_read of the same
 * request may run concurrently: the film is released only after it. */
int psy_server_wait(psy_server* server, uint64_t request_id, float* xyz, psy_request_stats* stats);
/* Incremental delivery: the number of leading pixels of request `id`
 * (row-major from the top row of its crop region, or of the image) that are
 * final, i.e. whose samples have all
 * been rendered, and a copy of final pixels [pixel_begin, pixel_end) as XYZ
 * floats into host memory.  Valid until psy_server_wait collects the
 * request; reading pixels that are not final fails with PSY_E_RANGE.  The
 * finished rows are what the reference streams as buckets to Blender
 * (src/renderer.rs:312-346). */
int psy_server_progress(psy_server* server, uint64_t id, uint64_t* pixels_done);
int psy_server_read(psy_server* server, uint64_t id, uint64_t pixel_begin, uint64_t pixel_end, float* xyz);

int psy_server_get_stats(psy_server* server, psy_server_stats* stats);

/* Test hook (no counterpart in the reference): the next `count` film
 * allocations of the server fail as if device memory were exhausted, so the
 * error path above can be exercised without exhausting the device. */
int psy_server_fail_film_allocs(psy_server* server, uint32_t count);

#ifdef __cplusplus
}
#endif

#endif /* PSY_RENDER_H */
