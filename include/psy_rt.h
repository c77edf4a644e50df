/*
 * psy_rt.h — C ABI of the MI355X (gfx950) ray-tracing path for psychopath.
 *
 * This is the drop-in boundary under the renderer.  It replaces
 *   Tracer::from_assembly(&Assembly) -> Tracer      (src/tracer.rs:22-32)
 *   Tracer::trace(&mut RayBatch) -> &[SurfaceIntersection]  (src/tracer.rs:34-37)
 *   Tracer::rays_traced() -> u64                    (src/tracer.rs:39-41)
 * together with the scene data those calls read: triangle meshes
 * (TriangleMesh::from_verts_and_indices, src/surface/triangle_mesh.rs:29-116)
 * and instances with time-sampled transforms (AssemblyBuilder::add_instance /
 * build, src/scene/assembly.rs:223-350).  The BVH4s are built on the host by
 * the reference's builder algorithm (src/accel/bvh_base.rs,
 * src/accel/objects_split.rs, src/accel/bvh4.rs:58-95,186-358) and then
 * repacked for the GPU.
 *
 * Conventions: every function returns PSY_OK (0) or a negative PSY_E* code;
 * psy_last_error() gives a message for the calling thread.  No exception
 * crosses the ABI.  Buffers are caller-owned.  A tracer owns one HIP stream
 * and is used by one host thread at a time (the reference keeps one Tracer
 * per render thread, src/renderer.rs:210).
 */
#ifndef PSY_RT_H
#define PSY_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSY_OK 0
#define PSY_E_INVALID (-1)      /* bad argument or handle */
#define PSY_E_RANGE (-2)        /* size beyond a documented limit */
#define PSY_E_DEVICE (-3)       /* HIP runtime error */
#define PSY_E_NOMEM (-4)        /* allocation failure */
#define PSY_E_STATE (-5)        /* call out of order */

/* Hit kinds, the variants of SurfaceIntersection (src/surface/mod.rs:72-80). */
#define PSY_HIT_MISS 0u
#define PSY_HIT_OCCLUDE 1u
#define PSY_HIT_SURFACE 2u

/* Ray flags (src/ray.rs:9-10). */
#define PSY_RAY_OCCLUSION 1u

/* Memory-space flags for psy_trace(). */
#define PSY_MEM_HOST 0u
#define PSY_MEM_DEVICE 1u

#define PSY_NONE 0xffffffffu

typedef struct psy_scene psy_scene;
typedef struct psy_tracer psy_tracer;

/* One triangle mesh (TriangleMesh::from_verts_and_indices,
 * src/surface/triangle_mesh.rs:29-116).  `verts` holds `time_samples`
 * consecutive arrays of `vert_count` xyz points (deformation motion blur);
 * `normals` is NULL or has the same layout; `indices` holds `tri_count`
 * triples in the file's winding order. */
typedef struct {
    uint32_t vert_count;
    uint32_t time_samples;
    const float* verts;
    const float* normals;
    uint32_t tri_count;
    const uint32_t* indices;
} psy_mesh_desc;

/* Colours (Color, src/color.rs:46-62): XYZ tristimulus (a, b, c = X, Y, Z),
 * or a blackbody spectrum of temperature `a` kelvin scaled by `b`, plain
 * (Blackbody) or normalised to roughly equal energy (Temperature). */
#define PSY_COLOR_XYZ 0u
#define PSY_COLOR_BLACKBODY 1u
#define PSY_COLOR_TEMPERATURE 2u
typedef struct {
    uint32_t kind;
    float a, b, c;
} psy_color;

/* Light objects of an assembly (RectangleLight::new,
 * src/light/rectangle_light.rs:31-48; SphereLight::new,
 * src/light/sphere_light.rs:34-48).  Both emit from both sides; the
 * rectangle lies in its local xy plane, centred.  Sizes and colours may be
 * time samples: a non-zero count replaces the value before it by that many
 * samples, lerped at each ray's time (lerp_slice, src/lerp.rs:20-35), each
 * list on its own; at most 64 per list, and the colours of one list must
 * share a representation (the reference cannot lerp between them).  The
 * light's bounds are one box per size sample and its energy, for light
 * selection, the mean over its colours. */
typedef struct {
    float width, height;
    psy_color color;
    uint32_t dims_count;      /* 0 = width, height above */
    const float* dims;        /* dims_count (width, height) pairs */
    uint32_t color_count;     /* 0 = color above */
    const psy_color* colors;
} psy_rect_light_desc;
typedef struct {
    float radius;
    psy_color color;
    uint32_t radius_count;    /* 0 = radius above */
    const float* radii;
    uint32_t color_count;     /* 0 = color above */
    const psy_color* colors;
} psy_sphere_light_desc;

/* Object kinds of an instance (Object, src/scene/assembly.rs:399-402, and
 * InstanceType::Assembly, :413-417). */
#define PSY_OBJECT_MESH 0u
#define PSY_OBJECT_RECT_LIGHT 1u
#define PSY_OBJECT_SPHERE_LIGHT 2u
#define PSY_OBJECT_ASSEMBLY 3u

/* One instance of an object (Instance, src/scene/assembly.rs:404-411).
 * `mesh` indexes meshes, rect_lights, sphere_lights or assemblies according
 * to `kind`.  `xform_count` transforms starting at `xform_start` in
 * psy_scene_desc.xforms are its time samples, from the enclosing assembly's
 * space to the object's (0 = no transform; a light instance of the root
 * assembly needs at least one).  Shader indices refer to the one shader
 * list the renderer takes, in every assembly. */
typedef struct {
    uint32_t mesh;
    uint32_t shader;      /* surface shader index, PSY_NONE = unassigned */
    uint32_t xform_start;
    uint32_t xform_count;
    uint32_t kind;        /* PSY_OBJECT_* */
} psy_instance_desc;

/* An assembly (Assembly, src/scene/assembly.rs:18-40): meshes, light
 * objects, sub-assemblies and the instances that place them.  Transforms are
 * 12 floats each, the arguments a..l of Transform::new_from_values
 * (sub_crates/math3d/src/transform.rs:29-47).  A sub-assembly is instanced
 * by its parent's instances of kind PSY_OBJECT_ASSEMBLY; its indices refer
 * to its own arrays.  At most 64 levels of nesting.
 *
 * Nested assemblies are traversed nested, as Tracer::trace_assembly
 * recurses (src/tracer.rs:87-150): every assembly keeps its own instance BVH
 * in its own space, a ray entering a sub-assembly instance moves into it by
 * the lerped transforms TransformStack::push composes along its path
 * (src/transform_stack.rs:33-60), and leaves it when that BVH is done.  Hit
 * records number the object instances of all paths depth first, in instance
 * order (so a single-level scene keeps its instance indices).  Light
 * selection keeps the reference's
 * hierarchy: one light tree per assembly over its lights and its light-
 * holding sub-assemblies (assembly.rs:297-338), selected recursively as
 * Assembly::sample_lights does (assembly.rs:46-141). */
typedef struct psy_scene_desc {
    uint32_t mesh_count;
    const psy_mesh_desc* meshes;
    uint32_t instance_count;
    const psy_instance_desc* instances;
    uint32_t xform_count;
    const float* xforms;
    uint32_t rect_light_count;
    const psy_rect_light_desc* rect_lights;
    uint32_t sphere_light_count;
    const psy_sphere_light_desc* sphere_lights;
    uint32_t assembly_count;
    const struct psy_scene_desc* assemblies;
} psy_scene_desc;

/* Rays as structure-of-arrays (RayBatch, src/ray.rs:24-44).  `flags` may be
 * NULL (all closest-hit).  `time` in [0,1]. */
typedef struct {
    const float *ox, *oy, *oz;
    const float *dx, *dy, *dz;
    const float* time;
    const float* max_t;
    const uint8_t* flags;
} psy_ray_soa;

/* One hit record, 32 bytes: the array-of-records form of the output, one
 * record per ray like the slice of SurfaceIntersection that Tracer::trace
 * returns (src/tracer.rs:34-37, src/surface/mod.rs:72-92).  Every field is
 * written: for a miss or an occlusion t and the barycentrics are 0 and the
 * ids PSY_NONE. */
typedef struct {
    float t, b0, b1, b2;
    uint32_t prim_id, inst_id;
    uint32_t kind; /* PSY_HIT_* */
    uint32_t reserved;
} psy_hit_record;

/* Hit output.  Barycentrics refer to the stored (winding-swapped) vertex
 * order (src/surface/triangle_mesh.rs:80): b0 to the first vertex, b1 to the
 * file's third, b2 to the file's second.  prim_id is the triangle's index in
 * the mesh description, inst_id the instance's index in
 * psy_scene_desc.instances.  On a light instance prim_id is the rectangle's
 * triangle (0: corners 1,2,3; 1: corners 3,4,1, rectangle_light.rs:288) or 0
 * for a sphere, whose barycentrics are 0.
 * When `records` is set the hits are written there (16-byte aligned, one
 * psy_hit_record per ray) and the arrays may be NULL; otherwise they go to
 * the arrays as structure-of-arrays, where fields other than `kind` are
 * written for PSY_HIT_SURFACE only. */
typedef struct {
    uint8_t* kind;
    float *t, *b0, *b1, *b2;
    uint32_t *prim_id, *inst_id;
    psy_hit_record* records;
} psy_hit_soa;

/* The geometry of one hit, 96 bytes: the remaining fields of the
 * reference's SurfaceIntersectionData (src/surface/mod.rs:82-92; `incoming`
 * is the ray's direction and `t` the hit record's), as the mesh and light
 * intersection code builds them (triangle_mesh.rs:264-317,
 * rectangle_light.rs:288-330, sphere_light.rs:283-330). */
typedef struct {
    float pos[3];          /* hit point, world space (triangle::surface_point) */
    float pos_err;         /* its error magnitude */
    float nor[3];          /* shading normal (interpolated vertex normals, else nor_g) */
    float nor_g[3];        /* geometric normal, cross(p0 - p1, p0 - p2), unnormalised */
    float local_space[12]; /* local_space as the 12 values a..l of Transform::new_from_values */
    float sample_pdf;      /* light hits: the light's pdf of sampling this point from the ray's origin; else 0 */
    float reserved;
} psy_hit_geometry;

/* Per-scene figures, including the record sizes used for the algorithmic-byte
 * roofline model. */
typedef struct {
    uint64_t node_count;        /* BVH4 internal nodes, all trees */
    uint64_t leaf_count;        /* BVH4 leaves, all trees */
    uint64_t motion_bounds;     /* extra time-sample BBox4 records */
    uint64_t tri_count;
    uint64_t device_bytes;      /* scene bytes resident in HBM */
    uint32_t node_bytes;        /* bytes of one node record */
    uint32_t motion_bounds_bytes; /* bytes of one time-sample BBox4 record */
    uint32_t tri_bytes;         /* bytes of one static triangle record */
    uint32_t max_depth;         /* deepest BVH4 (src/accel/bvh4.rs:87) */
    double build_seconds;       /* BVH build time (host and device builds) */
    uint32_t device_built_meshes; /* mesh BVHs the device builder built */
    double device_build_ms;     /* their device build time, transfers included */
    uint32_t device_light_trees;  /* light trees partitioned on the device */
} psy_scene_info;

/* Traversal counters of one psy_trace_counted() call. */
typedef struct {
    uint64_t node_visits;       /* internal-node box tests, summed over rays */
    uint64_t motion_node_visits;
    uint64_t tri_tests;         /* ray-triangle tests, summed over rays */
    uint64_t motion_tri_tests;  /* of which with deformation time samples */
    uint64_t instance_visits;   /* instance entries */
    uint64_t rays;
} psy_trace_counters;

/* Launch shape of the traversal kernel.  PSY_KERNEL_PERSISTENT_COOP is the
 * persistent kernel with a wave-cooperative node fetch: the nodes the lanes
 * of a wave stand on are loaded by the wave together, consecutive lanes
 * reading consecutive parts of one node, staged in LDS and read back, for
 * scenes without time samples (elsewhere it runs as the persistent kernel).
 * The persistent kernel keeps
 * `waves_per_cu` wave64 workgroups per compute unit resident, each taking
 * chunks of rays from a global counter; lanes whose ray finished are
 * refilled from the wave's chunk once `refill_min` lanes of a wave are idle; `stack_lds` stack entries per lane
 * live in LDS (8, 16 or 24) and the rest in private memory; `simd_waves`
 * is the register budget, as resident waves per SIMD (4: up to 128 VGPRs,
 * 8: at most 64); `leaf_min` 0 lets each lane descend until it reaches a
 * leaf or a node whose children all miss, 1..64 pops past such nodes inside
 * the descent and ends a wave's descent round once no lane stands on an
 * internal node or `leaf_min` lanes wait at a leaf.  The simple kernel runs
 * one lane per ray over a launch-wide grid.  No field changes the results. */
#define PSY_KERNEL_PERSISTENT 0u
#define PSY_KERNEL_SIMPLE 1u
#define PSY_KERNEL_PERSISTENT_COOP 2u
typedef struct {
    uint32_t kernel;
    uint32_t waves_per_cu;
    uint32_t refill_min;
    uint32_t stack_lds;
    uint32_t simd_waves;
    uint32_t leaf_min;
} psy_trace_tuning;

/* ---- errors ---------------------------------------------------------- */
const char* psy_last_error(void);
const char* psy_version(void);

/* ---- scene: AssemblyBuilder::build (src/scene/assembly.rs:288-350) ------
 * Builds every mesh BVH4 (<= 3 triangles per leaf, triangle_mesh.rs:17) and
 * the instance BVH4 (1 instance per leaf, assembly.rs:293) on the host.  The
 * repacked scene is uploaded to the current HIP device when the first tracer
 * is created, so scenes can be built and inspected without a GPU. */
int psy_scene_create(const psy_scene_desc* desc, psy_scene** out);
/* psy_scene_create with the mesh BVH4s of at least `min_tris` triangles (0:
 * every mesh) built on the current HIP device, level by level (binned SAH,
 * the Hoare partition's swaps from prefix counts, median splits near the
 * depth limit, bounds bottom-up, the BVH4 collapse): the host builder's
 * trees, bit for bit; light trees of at least `min_tris` lights (and two)
 * are partitioned on the device the same way (LightTree::from_objects,
 * light_tree.rs:211-290, SAH splits only).  A mesh or light list whose boxes
 * hold NaN or negative zero (their unions depend on order) or more than 64
 * time samples, or a HIP error, falls back to the host builder;
 * psy_scene_get_info counts what the device built. */
int psy_scene_create_on_device(const psy_scene_desc* desc, uint32_t min_tris, psy_scene** out);
int psy_scene_destroy(psy_scene* scene);
int psy_scene_get_info(const psy_scene* scene, psy_scene_info* info);

/* Canonical pre-order dump of one host-built BVH4 for parity checks against
 * the reference builder: mesh index, or -1 for the instance BVH.  Writes at
 * most `cap` u32 to `out` and the object order of the leaves (triangle or
 * instance indices) to `order` (up to `ocap`), and returns the dump length
 * in u32 (or a negative error). */
int64_t psy_scene_dump_bvh(const psy_scene* scene, int32_t mesh, uint32_t* out, uint64_t cap, uint32_t* order,
                           uint64_t ocap);

/* ---- tracer: Tracer::from_assembly (src/tracer.rs:22-32) ------------------
 * `max_batch` bounds the rays of one psy_trace() call when the rays are in
 * host memory (device staging is sized by it). */
int psy_tracer_create(const psy_scene* scene, uint64_t max_batch, psy_tracer** out);
int psy_tracer_destroy(psy_tracer* tracer);

/* Tracer::trace (src/tracer.rs:34-37): closest hit for ordinary rays, any hit
 * for rays flagged PSY_RAY_OCCLUSION, one record per ray.  `mem` says whether
 * `rays` and `hits` are host (PSY_MEM_HOST) or device (PSY_MEM_DEVICE)
 * pointers.  Synchronous: returns when the hits are written. */
int psy_trace(psy_tracer* tracer, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits, uint32_t mem);

/* As psy_trace with device pointers, but only enqueued on the tracer's
 * stream; psy_tracer_sync() waits.  The kernel time of the call is recorded
 * with HIP events on that stream. */
int psy_trace_async(psy_tracer* tracer, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits);
int psy_tracer_sync(psy_tracer* tracer);
/* Milliseconds of the last completed psy_trace / psy_trace_async kernel,
 * measured with HIP events on the tracer's stream. */
int psy_tracer_last_kernel_ms(psy_tracer* tracer, float* ms);

/* psy_trace with device pointers plus traversal counters (an instrumented
 * kernel: slower; the counts feed the roofline's algorithmic bytes and match
 * the meaning of ACCEL_NODE_RAY_TESTS, src/accel/mod.rs:22-25). */
int psy_trace_counted(psy_tracer* tracer, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits,
                      psy_trace_counters* counters);

/* Current launch shape of a tracer. */
int psy_tracer_get_tuning(const psy_tracer* tracer, psy_trace_tuning* tuning);
/* Sets the launch shape; values outside the documented ranges are rejected. */
int psy_tracer_set_tuning(psy_tracer* tracer, const psy_trace_tuning* tuning);
/* Autotuner: times every candidate launch shape (kernel, waves per CU,
 * refill threshold, LDS stack depth, register budget) on the given
 * device-resident rays,
 * keeps the fastest on the tracer and returns it in `best` (may be NULL).
 * `hits` is device scratch for the candidate runs.  Results do not depend on
 * the launch shape. */
int psy_tracer_autotune(psy_tracer* tracer, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits,
                        psy_trace_tuning* best);

/* The rest of each hit's SurfaceIntersectionData for rays traced by this
 * tracer's scene (device-resident rays, hits and output; synchronous on the
 * tracer's stream): `geometry[i]` is written for every PSY_HIT_SURFACE record
 * and zeroed for the others.  The reference computes these during
 * intersection (triangle_mesh.rs:264-317); here they are derived from the
 * hit record afterwards, with the same arithmetic, only for callers that
 * need them. */
int psy_finalize_hits(psy_tracer* tracer, const psy_ray_soa* rays, uint64_t n, const psy_hit_soa* hits,
                      psy_hit_geometry* geometry);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc after set-up).  psy_trace_async without a tracer and
 * without any allocation: device-resident rays and hits, enqueued on
 * `stream` (a hipStream_t, NULL for the default stream) and not waited for.
 * The scene must already be resident on the current device (creating a
 * tracer for it uploads it).  `tuning` may be NULL (the tracer's default
 * shape); the simple kernel is not offered.  The one scratch word the
 * persistent kernels need (their ray counter) lives in a fixed
 * PSY_NOALLOC_SCRATCH_BYTES device array of the library, one per device and
 * shared by every call on it: calls that overlap on one device race.  At
 * most PSY_NOALLOC_MAX_RAYS rays per call (PSY_E_RANGE beyond). */
#define PSY_NOALLOC_SCRATCH_BYTES 256u
#define PSY_NOALLOC_MAX_RAYS 0x7fffffffull
int psy_trace_noalloc(const psy_scene* scene, const psy_ray_soa* rays, uint64_t n, psy_hit_soa* hits,
                      const psy_trace_tuning* tuning, void* stream);

/* Tracer::rays_traced (src/tracer.rs:39-41). */
uint64_t psy_tracer_rays_traced(const psy_tracer* tracer);

/* ---- device memory helpers for callers without a GPU runtime of their own */
int psy_device_alloc(uint64_t bytes, void** ptr);
int psy_device_free(void* ptr);
int psy_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int psy_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
int psy_device_count(int* count);
/* A plain device-to-device copy kernel of `bytes` (multiple of 16), used to
 * calibrate HBM byte counters; returns its kernel time in ms. */
int psy_device_copy_kernel(void* dst, const void* src, uint64_t bytes, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* PSY_RT_H */
