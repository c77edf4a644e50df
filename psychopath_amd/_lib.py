"""ctypes binding of libpsy_rt.so (include/psy_rt.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
psychopath_amd/csrc``).  Loading fails loudly when it is missing: there is no
CPU fallback for the GPU path.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsy_rt.so")

PSY_OK = 0
PSY_E_INVALID, PSY_E_RANGE, PSY_E_DEVICE, PSY_E_NOMEM, PSY_E_STATE = -1, -2, -3, -4, -5
PSY_HIT_MISS = 0
PSY_HIT_OCCLUDE = 1
PSY_HIT_SURFACE = 2
PSY_RAY_OCCLUSION = 1
PSY_MEM_HOST = 0
PSY_MEM_DEVICE = 1
PSY_NONE = 0xFFFFFFFF

_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)


class MeshDesc(C.Structure):
    _fields_ = [
        ("vert_count", C.c_uint32),
        ("time_samples", C.c_uint32),
        ("verts", _fp),
        ("normals", _fp),
        ("tri_count", C.c_uint32),
        ("indices", _u32p),
    ]


PSY_COLOR_XYZ, PSY_COLOR_BLACKBODY, PSY_COLOR_TEMPERATURE = 0, 1, 2
PSY_OBJECT_MESH, PSY_OBJECT_RECT_LIGHT, PSY_OBJECT_SPHERE_LIGHT, PSY_OBJECT_ASSEMBLY = 0, 1, 2, 3
PSY_SHADER_LAMBERT, PSY_SHADER_GGX, PSY_SHADER_EMIT, PSY_SHADER_DIELECTRIC = 0, 1, 2, 3


class Color(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("a", C.c_float), ("b", C.c_float), ("c", C.c_float)]


class RectLightDesc(C.Structure):
    _fields_ = [("width", C.c_float), ("height", C.c_float), ("color", Color),
                ("dims_count", C.c_uint32), ("dims", _fp), ("color_count", C.c_uint32), ("colors", C.POINTER(Color))]


class SphereLightDesc(C.Structure):
    _fields_ = [("radius", C.c_float), ("color", Color),
                ("radius_count", C.c_uint32), ("radii", _fp), ("color_count", C.c_uint32), ("colors", C.POINTER(Color))]


class InstanceDesc(C.Structure):
    _fields_ = [
        ("mesh", C.c_uint32),
        ("shader", C.c_uint32),
        ("xform_start", C.c_uint32),
        ("xform_count", C.c_uint32),
        ("kind", C.c_uint32),
    ]


class SceneDesc(C.Structure):
    pass


SceneDesc._fields_ = [
    ("mesh_count", C.c_uint32),
    ("meshes", C.POINTER(MeshDesc)),
    ("instance_count", C.c_uint32),
    ("instances", C.POINTER(InstanceDesc)),
    ("xform_count", C.c_uint32),
    ("xforms", _fp),
    ("rect_light_count", C.c_uint32),
    ("rect_lights", C.POINTER(RectLightDesc)),
    ("sphere_light_count", C.c_uint32),
    ("sphere_lights", C.POINTER(SphereLightDesc)),
    ("assembly_count", C.c_uint32),
    ("assemblies", C.POINTER(SceneDesc)),
]


class ShaderDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("color", Color), ("roughness", C.c_float), ("fresnel", C.c_float),
                ("ior", C.c_float)]


class DistantLightDesc(C.Structure):
    _fields_ = [("radius", C.c_float), ("direction", C.c_float * 3), ("color", Color),
                ("radius_count", C.c_uint32), ("radii", _fp), ("direction_count", C.c_uint32), ("directions", _fp),
                ("color_count", C.c_uint32), ("colors", C.POINTER(Color))]


class WorldDesc(C.Structure):
    _fields_ = [
        ("shader_count", C.c_uint32),
        ("shaders", C.POINTER(ShaderDesc)),
        ("background", Color),
        ("distant_light_count", C.c_uint32),
        ("distant_lights", C.POINTER(DistantLightDesc)),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("xform_count", C.c_uint32),
        ("xforms", _fp),
        ("fov", C.c_float),
        ("aperture_radius", C.c_float),
        ("focus_distance", C.c_float),
        ("fov_count", C.c_uint32),
        ("fovs", _fp),
        ("aperture_count", C.c_uint32),
        ("aperture_radii", _fp),
        ("focus_count", C.c_uint32),
        ("focus_distances", _fp),
    ]


class RenderSettings(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("seed", C.c_uint32),
        ("crop_x", C.c_uint32),
        ("crop_y", C.c_uint32),
        ("crop_width", C.c_uint32),
        ("crop_height", C.c_uint32),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("paths", C.c_uint64),
        ("rays", C.c_uint64),
        ("waves", C.c_uint64),
        ("seconds", C.c_double),
        ("trace_ms", C.c_double),
    ]


class RaySoA(C.Structure):
    _fields_ = [
        ("ox", C.c_void_p), ("oy", C.c_void_p), ("oz", C.c_void_p),
        ("dx", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p),
        ("time", C.c_void_p), ("max_t", C.c_void_p), ("flags", C.c_void_p),
    ]


class HitSoA(C.Structure):
    _fields_ = [
        ("kind", C.c_void_p),
        ("t", C.c_void_p), ("b0", C.c_void_p), ("b1", C.c_void_p), ("b2", C.c_void_p),
        ("prim_id", C.c_void_p), ("inst_id", C.c_void_p),
        ("records", C.c_void_p),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("node_count", C.c_uint64),
        ("leaf_count", C.c_uint64),
        ("motion_bounds", C.c_uint64),
        ("tri_count", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("node_bytes", C.c_uint32),
        ("motion_bounds_bytes", C.c_uint32),
        ("tri_bytes", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("build_seconds", C.c_double),
        ("device_built_meshes", C.c_uint32),
        ("device_build_ms", C.c_double),
        ("device_light_trees", C.c_uint32),
    ]


class RenderProfile(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("camera_ms", "trace_ms", "shadow_trace_ms", "shadow_ms", "sort_ms",
                                           "shade_ms", "film_ms")] + \
               [(f, C.c_uint64) for f in ("paths", "closest_rays", "shadow_rays", "closest_node_visits",
                                           "closest_tri_tests", "shadow_node_visits", "shadow_tri_tests",
                                           "motion_node_visits", "motion_tri_tests", "instance_visits")]


class TraceCounters(C.Structure):
    _fields_ = [
        ("node_visits", C.c_uint64),
        ("motion_node_visits", C.c_uint64),
        ("tri_tests", C.c_uint64),
        ("motion_tri_tests", C.c_uint64),
        ("instance_visits", C.c_uint64),
        ("rays", C.c_uint64),
    ]


class TraceTuning(C.Structure):
    _fields_ = [
        ("kernel", C.c_uint32),
        ("waves_per_cu", C.c_uint32),
        ("refill_min", C.c_uint32),
        ("stack_lds", C.c_uint32),
        ("simd_waves", C.c_uint32),
        ("leaf_min", C.c_uint32),
    ]


PSY_KERNEL_PERSISTENT = 0
PSY_KERNEL_SIMPLE = 1
PSY_KERNEL_PERSISTENT_COOP = 2
PSY_NOALLOC_SCRATCH_BYTES = 256
PSY_NOALLOC_MAX_RAYS = 0x7FFFFFFF

# Every symbol the header declares, with its ctypes signature.
_SIGNATURES = {
    "psy_last_error": (C.c_char_p, []),
    "psy_version": (C.c_char_p, []),
    "psy_scene_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "psy_scene_create_on_device": (C.c_int, [C.POINTER(SceneDesc), C.c_uint32, C.POINTER(C.c_void_p)]),
    "psy_scene_destroy": (C.c_int, [C.c_void_p]),
    "psy_scene_get_info": (C.c_int, [C.c_void_p, C.POINTER(SceneInfo)]),
    "psy_scene_dump_bvh": (C.c_int64, [C.c_void_p, C.c_int32, _u32p, C.c_uint64, _u32p, C.c_uint64]),
    "psy_tracer_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "psy_tracer_destroy": (C.c_int, [C.c_void_p]),
    "psy_trace": (C.c_int, [C.c_void_p, C.POINTER(RaySoA), C.c_uint64, C.POINTER(HitSoA), C.c_uint32]),
    "psy_trace_async": (C.c_int, [C.c_void_p, C.POINTER(RaySoA), C.c_uint64, C.POINTER(HitSoA)]),
    "psy_trace_noalloc": (C.c_int, [C.c_void_p, C.POINTER(RaySoA), C.c_uint64, C.POINTER(HitSoA),
                                    C.POINTER(TraceTuning), C.c_void_p]),
    "psy_finalize_hits": (C.c_int, [C.c_void_p, C.POINTER(RaySoA), C.c_uint64, C.POINTER(HitSoA), C.c_void_p]),
    "psy_tracer_sync": (C.c_int, [C.c_void_p]),
    "psy_tracer_last_kernel_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "psy_trace_counted": (C.c_int, [C.c_void_p, C.POINTER(RaySoA), C.c_uint64, C.POINTER(HitSoA),
                                    C.POINTER(TraceCounters)]),
    "psy_tracer_rays_traced": (C.c_uint64, [C.c_void_p]),
    "psy_tracer_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(TraceTuning)]),
    "psy_tracer_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(TraceTuning)]),
    "psy_tracer_autotune": (C.c_int, [C.c_void_p, C.POINTER(RaySoA), C.c_uint64, C.POINTER(HitSoA),
                                      C.POINTER(TraceTuning)]),
    "psy_device_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "psy_device_free": (C.c_int, [C.c_void_p]),
    "psy_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "psy_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "psy_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "psy_device_copy_kernel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_float)]),
    # include/psy_render.h
    "psy_renderer_create": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_uint64, C.POINTER(C.c_void_p)]),
    "psy_renderer_destroy": (C.c_int, [C.c_void_p]),
    "psy_render": (C.c_int, [C.c_void_p, C.POINTER(CameraDesc), C.POINTER(RenderSettings), C.c_void_p, C.c_uint32,
                             C.POINTER(RenderStats)]),
    "psy_renderer_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(TraceTuning)]),
    "psy_renderer_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(TraceTuning)]),
    "psy_renderer_set_options": (C.c_int, [C.c_void_p, C.c_uint32]),
    "psy_renderer_get_profile": (C.c_int, [C.c_void_p, C.POINTER(RenderProfile)]),
    "psy_renderer_autotune": (C.c_int, [C.c_void_p, C.POINTER(CameraDesc), C.POINTER(RenderSettings),
                                        C.POINTER(TraceTuning), C.POINTER(C.c_uint32)]),
}

PSY_RENDER_NO_GRAPHS = 1
PSY_RENDER_NO_SORT = 2
PSY_RENDER_PLAIN = 4
PSY_RENDER_PROFILE = 16
PSY_RENDER_COUNT = 32
PSY_RENDER_SHADE_WAVES_SHIFT = 8


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_active", C.c_uint32)]


class RequestStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("waves", C.c_uint64), ("queued_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("waves", C.c_uint64), ("paths", C.c_uint64), ("queued", C.c_uint64), ("fixed_bytes", C.c_uint64),
                ("film_bytes", C.c_uint64)]


_SIGNATURES.update({
    "psy_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(ServerConfig),
                                    C.POINTER(C.c_void_p)]),
    "psy_server_destroy": (C.c_int, [C.c_void_p]),
    "psy_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(CameraDesc), C.POINTER(RenderSettings),
                                    C.POINTER(C.c_uint64)]),
    "psy_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(RequestStats)]),
    "psy_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "psy_server_progress": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "psy_server_read": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "psy_server_fail_film_allocs": (C.c_int, [C.c_void_p, C.c_uint32]),
})


class FileScene(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("output_path", C.c_char_p),
        ("scene", SceneDesc),
        ("world", WorldDesc),
        ("camera", CameraDesc),
        ("settings", RenderSettings),
    ]


PSY_E_PARSE = -6

_SIGNATURES.update({
    # include/psy_file.h
    "psy_file_parse": (C.c_int, [C.c_char_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "psy_file_destroy": (C.c_int, [C.c_void_p]),
    "psy_file_scene_count": (C.c_uint32, [C.c_void_p]),
    "psy_file_get_scene": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(FileScene)]),
    "psy_file_dump_tree": (C.c_int64, [C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint64]),
    # include/psy_image.h
    "psy_image_encode": (C.c_int64, [_fp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64]),
    "psy_image_bucket_base64": (C.c_int64, [_fp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_char_p, C.c_uint64]),
})

PSY_IMAGE_PPM_ASCII, PSY_IMAGE_PPM, PSY_IMAGE_PNG, PSY_IMAGE_EXR = 0, 1, 2, 3


class PsyError(RuntimeError):
    """A psy_rt call returned an error code."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Returns the loaded library, loading it on first use."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the GPU path has no fallback)")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(fn: str, rc: int) -> None:
    if rc != PSY_OK:
        msg = lib().psy_last_error().decode(errors="replace")
        raise PsyError(fn, rc, msg)


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)
