"""Host-side mirror of the reference's renderer interface over the C ABI
(include/psy_render.h).

    Renderer       ~ Renderer::render              (src/renderer.rs:32-192)
    World          ~ World + the scene's shaders   (src/scene/world.rs:4-8,
                                                    src/shading/mod.rs:28-61)
    Camera         ~ Camera::new                   (src/camera.rs:22-72)

`build(render_scene)` turns a psychopath_amd.scenes.RenderScene into a
device scene plus renderer.  Images are XYZ float32 arrays [H, W, 3], row 0
at the top.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .tracer import Assembly, Mesh, RectLight, SphereLight, color_tuple, colour_samples, time_samples


def _color(c) -> _lib.Color:
    return _lib.Color(*color_tuple(c))


class Renderer:
    """A device-resident scene with its shaders, world and path state."""

    def __init__(self, scene, shaders, background=(0, 0.0, 0.0, 0.0), distant_lights=(), max_paths: int = 0):
        self.scene = scene  # keeps the scene alive
        sh = (_lib.ShaderDesc * max(1, len(shaders)))()
        for i, (kind, col, rough, fres) in enumerate(shaders):
            # (kind, colour, roughness, fresnel): the last is the index of
            # refraction for a dielectric.
            if int(kind) == _lib.PSY_SHADER_DIELECTRIC:
                sh[i] = _lib.ShaderDesc(int(kind), _color(col), float(rough), 0.0, float(fres))
            else:
                sh[i] = _lib.ShaderDesc(int(kind), _color(col), float(rough), float(fres), 0.0)
        dl = (_lib.DistantLightDesc * max(1, len(distant_lights)))()
        keep = []
        for i, (radius, direction, col) in enumerate(distant_lights):
            # Numbers and tuples are one sample; lists are time samples.
            r0, nr, ra = time_samples(radius, 1)
            d0, nd, da = time_samples(direction, 3)
            c0, nc, ca = colour_samples(col)
            dl[i] = _lib.DistantLightDesc(float(r0[0]), (C.c_float * 3)(*[float(x) for x in d0]), c0, nr,
                                          ra.ctypes.data_as(_lib._fp), nd, da.ctypes.data_as(_lib._fp), nc, ca)
            keep += [ra, da, ca]
        world = _lib.WorldDesc(len(shaders), sh, _color(background), len(distant_lights), dl)
        h = C.c_void_p()
        check("psy_renderer_create", lib().psy_renderer_create(scene.handle, C.byref(world), int(max_paths),
                                                                C.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            _lib._lib.psy_renderer_destroy(h)
            self.handle = None

    def set_options(self, graphs: bool = True, sort: bool = True, plain: bool = False, shade_waves: int = 0,
                    profile: bool = False, count: bool = False) -> None:
        """graphs/sort: wavefront execution choices; plain: one launch per
        render over the whole frame; shade_waves (1-4, 0 keeps the current):
        the shading kernels' build by minimum waves per SIMD; profile:
        per-stage HIP-event times and ray counts; count: instrumented
        traversal counters (see profile()).  None of them changes the
        image."""
        flags = ((0 if graphs else _lib.PSY_RENDER_NO_GRAPHS) | (0 if sort else _lib.PSY_RENDER_NO_SORT) |
                 (_lib.PSY_RENDER_PLAIN if plain else 0) | (int(shade_waves) << _lib.PSY_RENDER_SHADE_WAVES_SHIFT) |
                 (_lib.PSY_RENDER_PROFILE if profile else 0) | (_lib.PSY_RENDER_COUNT if count else 0))
        check("psy_renderer_set_options", lib().psy_renderer_set_options(self.handle, flags))

    def profile(self) -> dict:
        """psy_renderer_get_profile: stage times (ms), paths, rays and
        traversal counters tallied since the last call."""
        p = _lib.RenderProfile()
        check("psy_renderer_get_profile", lib().psy_renderer_get_profile(self.handle, C.byref(p)))
        return {f: getattr(p, f) for f, _ in _lib.RenderProfile._fields_}

    def autotune(self, camera: dict, trial: dict) -> dict:
        """Picks the traversal launch shape and material sorting by timing
        `trial` renders; returns the choice."""
        cam, _xf = self._camera(camera)
        st = self._settings(trial)
        t = _lib.TraceTuning()
        flags = C.c_uint32()
        check("psy_renderer_autotune", lib().psy_renderer_autotune(self.handle, C.byref(cam), C.byref(st),
                                                                    C.byref(t), C.byref(flags)))
        out = {f: getattr(t, f) for f, _ in _lib.TraceTuning._fields_}
        out["sort"] = not (flags.value & _lib.PSY_RENDER_NO_SORT)
        out["plain"] = bool(flags.value & _lib.PSY_RENDER_PLAIN)
        out["shade_waves"] = (flags.value >> _lib.PSY_RENDER_SHADE_WAVES_SHIFT) & 7
        return out

    def tuning(self) -> dict:
        t = _lib.TraceTuning()
        check("psy_renderer_get_tuning", lib().psy_renderer_get_tuning(self.handle, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.TraceTuning._fields_}

    def set_tuning(self, **kw) -> None:
        t = _lib.TraceTuning(**{**self.tuning(), **kw})
        check("psy_renderer_set_tuning", lib().psy_renderer_set_tuning(self.handle, C.byref(t)))

    @staticmethod
    def _camera(camera: dict):
        """The camera descriptor and the arrays it points into.  `fov`,
        `aperture` and `focus` are numbers or lists of time samples."""
        xf = np.ascontiguousarray(camera["xforms"], np.float32).reshape(-1, 12)
        lens = [np.ascontiguousarray(np.atleast_1d(camera.get(k, d)), np.float32).ravel()
                for k, d in (("fov", None), ("aperture", 0.0), ("focus", 1.0))]
        cam = _lib.CameraDesc(xf.shape[0], xf.ctypes.data_as(_lib._fp), float(lens[0][0]), float(lens[1][0]),
                              float(lens[2][0]))
        for name, a in zip(("fov", "aperture", "focus"), lens):
            if a.size > 1:
                field = {"fov": "fovs", "aperture": "aperture_radii", "focus": "focus_distances"}[name]
                setattr(cam, name + "_count", a.size)
                setattr(cam, field, a.ctypes.data_as(_lib._fp))
        return cam, (xf, lens)

    @staticmethod
    def _settings(settings: dict):
        """`crop` (optional): (x, y, width, height) of the region to render."""
        cx, cy, cw, ch = settings.get("crop") or (0, 0, 0, 0)
        return _lib.RenderSettings(int(settings["width"]), int(settings["height"]), int(settings["spp"]),
                                   int(settings.get("max_bounces", 2)), int(settings.get("seed", 0)), int(cx), int(cy),
                                   int(cw), int(ch))

    def render(self, camera: dict, settings: dict, device_out=None) -> tuple[np.ndarray | None, dict]:
        """Renders one image.  `camera`: xforms [N,12] (camera to world), fov
        (radians), aperture, focus (each a number or a list of time
        samples).  `settings`: width, height, spp,
        max_bounces, seed.  With `device_out` (a device pointer of
        width*height*3 floats) the image stays on the device."""
        cam, _xf = self._camera(camera)
        st = self._settings(settings)
        stats = _lib.RenderStats()
        if device_out is None:
            img = np.zeros((st.height, st.width, 3), np.float32)
            ptr, mem = img.ctypes.data, _lib.PSY_MEM_HOST
        else:
            img, ptr, mem = None, int(device_out), _lib.PSY_MEM_DEVICE
        check("psy_render", lib().psy_render(self.handle, C.byref(cam), C.byref(st), C.c_void_p(ptr), mem,
                                             C.byref(stats)))
        return img, {f: getattr(stats, f) for f, _ in _lib.RenderStats._fields_}


def assembly_of(rs) -> Assembly:
    """The tracer Assembly of a RenderScene (or of one of its optional
    `sub_assemblies`: objects with meshes, rect_lights, sphere_lights,
    instances and sub_assemblies, instances of kind "assembly" naming a
    sub-assembly by index)."""
    asm = Assembly()
    for m in rs.meshes:
        asm.add_mesh(Mesh(m[0], m[1], m[2] if len(m) > 2 else None))
    for (w, h, col) in getattr(rs, "rect_lights", []):
        asm.add_rect_light(RectLight(w, h, col))
    for (r, col) in getattr(rs, "sphere_lights", []):
        asm.add_sphere_light(SphereLight(r, col))
    for sub in getattr(rs, "sub_assemblies", []):
        asm.add_assembly(assembly_of(sub))
    for inst in rs.instances:
        kind = {"mesh": _lib.PSY_OBJECT_MESH, "rect": _lib.PSY_OBJECT_RECT_LIGHT,
                "sphere": _lib.PSY_OBJECT_SPHERE_LIGHT, "assembly": _lib.PSY_OBJECT_ASSEMBLY}[inst["kind"]]
        shader = _lib.PSY_NONE if inst.get("shader") is None else int(inst["shader"])
        asm.add_instance(inst["index"], inst.get("xforms"), shader, kind)
    return asm


def build(rs, max_paths: int = 0) -> Renderer:
    """Device scene + renderer for a psychopath_amd.scenes.RenderScene."""
    scene = assembly_of(rs).build()
    return Renderer(scene, rs.shaders, rs.background, rs.distant_lights, max_paths)


class Server:
    """Render request server over several renderers (psy_server_*):
    continuous batching of requests into shared waves, HBM-budget
    admission, one stream per scene."""

    def __init__(self, renderers, hbm_budget: int = 0, max_active: int = 0):
        self.renderers = list(renderers)  # keeps them alive
        arr = (C.c_void_p * len(self.renderers))(*[r.handle.value for r in self.renderers])
        cfg = _lib.ServerConfig(int(hbm_budget), int(max_active))
        h = C.c_void_p()
        check("psy_server_create", lib().psy_server_create(arr, len(self.renderers), C.byref(cfg), C.byref(h)))
        self.handle = h
        self._settings = {}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().psy_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if _lib._lib is not None:
            self.close()

    def submit(self, renderer: int, camera: dict, settings: dict) -> int:
        cam, _xf = Renderer._camera(camera)
        st = Renderer._settings(settings)
        rid = C.c_uint64()
        check("psy_server_submit", lib().psy_server_submit(self.handle, int(renderer), C.byref(cam), C.byref(st),
                                                            C.byref(rid)))
        rw, rh = (st.crop_width, st.crop_height) if st.crop_width else (st.width, st.height)
        self._settings[rid.value] = (st.height, st.width, rh, rw)
        return rid.value

    def wait(self, request_id: int, keep_image: bool = True) -> tuple[np.ndarray | None, dict]:
        h, w, _, _ = self._settings.pop(request_id)
        img = np.zeros((h, w, 3), np.float32) if keep_image else None
        rs = _lib.RequestStats()
        check("psy_server_wait", lib().psy_server_wait(self.handle, int(request_id),
                                                        None if img is None else img.ctypes.data_as(C.c_void_p),
                                                        C.byref(rs)))
        return img, {f: getattr(rs, f) for f, _ in _lib.RequestStats._fields_}

    def progress(self, request_id: int) -> int:
        """Leading pixels of the request (row-major from the top) that are
        final."""
        n = C.c_uint64()
        check("psy_server_progress", lib().psy_server_progress(self.handle, int(request_id), C.byref(n)))
        return n.value

    def read(self, request_id: int, begin: int, end: int) -> np.ndarray:
        """Final pixels [begin, end) as XYZ [end - begin, 3]."""
        out = np.zeros((max(0, end - begin), 3), np.float32)
        check("psy_server_read", lib().psy_server_read(self.handle, int(request_id), int(begin), int(end),
                                                        out.ctypes.data_as(C.c_void_p)))
        return out

    def stream(self, request_id: int, poll_s: float = 0.002):
        """Yields (row_begin, rows) as whole rows of the request (of its crop
        region, if cropped) become final, then collects it; rows are XYZ
        [n, width, 3].  The rows are what the reference streams to Blender
        bucket by bucket."""
        import time
        _, _, h, w = self._settings[request_id]
        sent = 0
        while sent < h:
            done_rows = min(h, self.progress(request_id) // w)
            if done_rows > sent:
                yield sent, self.read(request_id, sent * w, done_rows * w).reshape(done_rows - sent, w, 3)
                sent = done_rows
            else:
                time.sleep(poll_s)
        self.wait(request_id, keep_image=False)

    def fail_film_allocs(self, count: int) -> None:
        """Test hook: the next `count` film allocations fail as if device
        memory were exhausted."""
        check("psy_server_fail_film_allocs", lib().psy_server_fail_film_allocs(self.handle, int(count)))

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check("psy_server_get_stats", lib().psy_server_get_stats(self.handle, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.ServerStats._fields_}


def main(argv=None) -> int:
    """Renders one of the built-in workload scenes and prints per-render
    timing, the way the reference's command line renders a scene file
    (src/main.rs: spp and output options); the image is written as a
    float32 XYZ .npy array."""
    import argparse
    import time

    from . import scenes

    makers = {
        "cornell": scenes.cornell_box,
        "sphere": scenes.sphere_scene,
        "motion": scenes.motion_instances,
        "terrain": scenes.terrain_scene,
        "large": scenes.large_scene,
    }
    ap = argparse.ArgumentParser(prog="python -m psychopath_amd.render")
    ap.add_argument("scene", choices=sorted(makers))
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("-s", "--spp", type=int)
    ap.add_argument("--max-bounces", type=int)
    ap.add_argument("--seed", type=int)
    ap.add_argument("--renders", type=int, default=1, help="renders after one untimed warm-up render")
    ap.add_argument("--plain", action="store_true", help="one launch per render over the whole frame")
    ap.add_argument("--autotune", action="store_true", help="pick the launch shape on a 1/16-size trial render")
    ap.add_argument("--shade-waves", type=int, default=0, help="shading kernels' build by waves per SIMD (1-4)")
    ap.add_argument("-o", "--output", help="write the last render: .png, .ppm or .exr as the reference writes "
                                           "them, or .npy for the raw XYZ film")
    ap.add_argument("--blender", action="store_true",
                    help="print the last render as the reference's serialized bucket output")
    a = ap.parse_args(argv)
    t0 = time.time()
    sc = makers[a.scene]()
    st = dict(sc.settings)
    for key, val in (("width", a.width), ("height", a.height), ("spp", a.spp), ("max_bounces", a.max_bounces),
                     ("seed", a.seed)):
        if val is not None:
            st[key] = val
    r = build(sc)
    print(f"{a.scene}: scene ready in {time.time() - t0:.1f}s", flush=True)
    if a.autotune:
        trial = dict(st, width=max(1, st["width"] // 4), height=max(1, st["height"] // 4),
                     spp=max(1, st["spp"] // 16))
        print(f"autotune: {r.autotune(sc.camera, trial)}", flush=True)
    if a.plain or a.shade_waves:
        r.set_options(plain=a.plain, shade_waves=a.shade_waves)
    img, stats = r.render(sc.camera, st)
    for _ in range(a.renders):
        t = time.perf_counter()
        img, stats = r.render(sc.camera, st)
        s = time.perf_counter() - t
        print(f"{a.scene} {st['width']}x{st['height']} {st['spp']} spp depth {st['max_bounces']}"
              f"{' plain' if a.plain else ''}: {s * 1e3:.2f} ms, {stats['paths'] / s / 1e6:.1f} M samples/s, "
              f"{stats['waves']} waves", flush=True)
    if a.output:
        if a.output.lower().endswith(".npy"):
            np.save(a.output, img)
        else:
            from . import film
            film.write(a.output, img)
    if a.blender:
        import sys

        from . import film
        for block in film.blender_buckets(img):
            sys.stdout.write(block)
        sys.stdout.flush()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
