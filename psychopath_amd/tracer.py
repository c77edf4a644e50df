"""Host-side mirror of the reference's tracing interface over the C ABI.

    Assembly       ~ AssemblyBuilder / Assembly   (src/scene/assembly.rs:150-350)
    Tracer         ~ Tracer                       (src/tracer.rs:15-42)
    RayBatch       ~ RayBatch                     (src/ray.rs:24-195), SoA numpy
    Hits           ~ [SurfaceIntersection]        (src/surface/mod.rs:70-92)

Rays and hits are numpy structure-of-arrays; `DeviceRays` / `DeviceHits` keep
them resident in HBM for repeated traces.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib

RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "time", "max_t")
HIT_FLOATS = ("t", "b0", "b1", "b2")
HIT_U32 = ("prim_id", "inst_id")
# psy_hit_record: one 32-byte record per ray.
HIT_RECORD = np.dtype([("t", "<f4"), ("b0", "<f4"), ("b1", "<f4"), ("b2", "<f4"), ("prim_id", "<u4"),
                       ("inst_id", "<u4"), ("kind", "<u4"), ("reserved", "<u4")])
assert HIT_RECORD.itemsize == 32
# psy_hit_geometry: the rest of SurfaceIntersectionData, 96 bytes per ray.
HIT_GEOMETRY = np.dtype([("pos", "<f4", 3), ("pos_err", "<f4"), ("nor", "<f4", 3), ("nor_g", "<f4", 3),
                         ("local_space", "<f4", 12), ("sample_pdf", "<f4"), ("reserved", "<f4")])
assert HIT_GEOMETRY.itemsize == 96


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


@dataclass
class Mesh:
    """A triangle mesh: verts [time_samples, V, 3], indices [T, 3],
    optional normals like verts (TriangleMesh, triangle_mesh.rs:29-116)."""
    verts: np.ndarray
    indices: np.ndarray
    normals: np.ndarray | None = None

    def __post_init__(self):
        v = _f32(self.verts)
        if v.ndim == 2:
            v = v[None]
        if v.ndim != 3 or v.shape[2] != 3:
            raise ValueError("verts must be [V,3] or [S,V,3]")
        self.verts = v
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)
        if self.normals is not None:
            n = _f32(self.normals)
            if n.ndim == 2:
                n = n[None]
            if n.shape != v.shape:
                raise ValueError("normals must match verts")
            self.normals = n


def color_tuple(c) -> tuple:
    """A colour as (kind, a, b, c): kind 0 = XYZ, 1 = blackbody (temperature,
    factor), 2 = normalised blackbody (Color, src/color.rs:46-62)."""
    k, a, b, cc = c
    return (int(k), float(a), float(b), float(cc))


def time_samples(v, width: int):
    """A light parameter as (first value, count, array): a number or tuple
    is one sample (count 0, as the descriptors take it), a list holds time
    samples of `width` floats each."""
    a = np.ascontiguousarray(np.asarray(v, np.float32).reshape(-1, width))
    return a[0], (a.shape[0] if a.shape[0] > 1 else 0), a


def colour_samples(c):
    """A colour or a list of colour time samples as (first, count,
    Color array or None)."""
    cs = [color_tuple(x) for x in c] if isinstance(c, list) else [color_tuple(c)]
    if len(cs) == 1:
        return _lib.Color(*cs[0]), 0, None
    return _lib.Color(*cs[0]), len(cs), (_lib.Color * len(cs))(*[_lib.Color(*x) for x in cs])


@dataclass
class RectLight:
    """RectangleLight (src/light/rectangle_light.rs:24-48).  `width` and
    `height` may be equal-length lists of time samples and `color` a list of
    colours."""
    width: float
    height: float
    color: tuple


@dataclass
class SphereLight:
    """SphereLight (src/light/sphere_light.rs:27-48).  `radius` and `color`
    may be lists of time samples."""
    radius: float
    color: tuple


@dataclass
class Instance:
    """An instance of object `mesh` (a mesh, or a light of kind `kind`) with
    world-to-object transforms, each the 12 values a..l of
    Transform::new_from_values (math3d transform.rs:29-47)."""
    mesh: int
    xforms: np.ndarray | None = None
    shader: int = _lib.PSY_NONE
    kind: int = _lib.PSY_OBJECT_MESH


@dataclass
class Assembly:
    """An assembly of meshes, light objects, sub-assemblies and the
    instances that place them (AssemblyBuilder::add_object / add_assembly /
    add_instance, assembly.rs:197-282)."""
    meshes: list = field(default_factory=list)
    instances: list = field(default_factory=list)
    rect_lights: list = field(default_factory=list)
    sphere_lights: list = field(default_factory=list)
    assemblies: list = field(default_factory=list)

    def add_assembly(self, sub: "Assembly") -> int:
        """A sub-assembly, placed by instances of kind PSY_OBJECT_ASSEMBLY
        with its index."""
        self.assemblies.append(sub)
        return len(self.assemblies) - 1

    def add_mesh(self, mesh: Mesh) -> int:
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_rect_light(self, light: RectLight) -> int:
        self.rect_lights.append(light)
        return len(self.rect_lights) - 1

    def add_sphere_light(self, light: SphereLight) -> int:
        self.sphere_lights.append(light)
        return len(self.sphere_lights) - 1

    def add_instance(self, mesh: int, xforms=None, shader: int = _lib.PSY_NONE,
                     kind: int = _lib.PSY_OBJECT_MESH) -> int:
        if xforms is not None:
            xforms = _f32(xforms).reshape(-1, 12)
            if len(xforms) == 0:
                xforms = None
        self.instances.append(Instance(mesh, xforms, shader, kind))
        return len(self.instances) - 1

    def build(self, device: bool = False, device_min_tris: int = 0) -> "Scene":
        """The scene with its BVHs built on the host, or (device=True) its
        mesh BVHs of at least `device_min_tris` triangles built on the
        current GPU (psy_scene_create_on_device): the same trees."""
        return Scene(self, device=device, device_min_tris=device_min_tris)


def _scene_desc(asm: Assembly, keep: list) -> _lib.SceneDesc:
    """psy_scene_desc of `asm` and its sub-assemblies; `keep` holds the
    arrays the descriptors point into."""
    mds = (_lib.MeshDesc * max(1, len(asm.meshes)))()
    for i, m in enumerate(asm.meshes):
        v = m.verts
        idx = m.indices
        keep += [v, idx]
        mds[i].vert_count = v.shape[1]
        mds[i].time_samples = v.shape[0]
        mds[i].verts = v.ctypes.data_as(_lib._fp)
        mds[i].normals = None
        if m.normals is not None:
            keep.append(m.normals)
            mds[i].normals = m.normals.ctypes.data_as(_lib._fp)
        mds[i].tri_count = idx.shape[0]
        mds[i].indices = idx.ctypes.data_as(_lib._u32p)
    xf_list = []
    ids = (_lib.InstanceDesc * max(1, len(asm.instances)))()
    nxf = 0
    for i, inst in enumerate(asm.instances):
        ids[i].mesh = inst.mesh
        ids[i].shader = inst.shader
        ids[i].kind = inst.kind
        ids[i].xform_start = nxf
        cnt = 0 if inst.xforms is None else len(inst.xforms)
        ids[i].xform_count = cnt
        if cnt:
            xf_list.append(inst.xforms)
        nxf += cnt
    xfa = _f32(np.concatenate(xf_list)) if xf_list else np.zeros((1, 12), np.float32)
    keep.append(xfa)
    rls = (_lib.RectLightDesc * max(1, len(asm.rect_lights)))()
    for i, l in enumerate(asm.rect_lights):
        # A number broadcasts against the other list (Dimensions leaves
        # always pair them in the file).
        wh = np.stack(np.broadcast_arrays(np.atleast_1d(np.asarray(l.width, np.float32)),
                                          np.atleast_1d(np.asarray(l.height, np.float32))), axis=1)
        d0, nd, da = time_samples(wh, 2)
        c0, nc, ca = colour_samples(l.color)
        rls[i] = _lib.RectLightDesc(float(d0[0]), float(d0[1]), c0, nd, da.ctypes.data_as(_lib._fp), nc, ca)
        keep += [da, ca]
    sls = (_lib.SphereLightDesc * max(1, len(asm.sphere_lights)))()
    for i, l in enumerate(asm.sphere_lights):
        r0, nr, ra = time_samples(l.radius, 1)
        c0, nc, ca = colour_samples(l.color)
        sls[i] = _lib.SphereLightDesc(float(r0[0]), c0, nr, ra.ctypes.data_as(_lib._fp), nc, ca)
        keep += [ra, ca]
    keep += [rls, sls]
    subs = (_lib.SceneDesc * max(1, len(asm.assemblies)))(*[_scene_desc(a, keep) for a in asm.assemblies])
    keep.append(subs)
    return _lib.SceneDesc(len(asm.meshes), mds, len(asm.instances), ids, nxf, xfa.ctypes.data_as(_lib._fp),
                          len(asm.rect_lights), rls, len(asm.sphere_lights), sls, len(asm.assemblies), subs)


class Scene:
    """The assembly uploaded to the GPU (psy_scene_create)."""

    def __init__(self, asm: Assembly, device: bool = False, device_min_tris: int = 0):
        self._keep = []
        desc = _scene_desc(asm, self._keep)
        h = C.c_void_p()
        if device:
            check("psy_scene_create_on_device",
                  lib().psy_scene_create_on_device(C.byref(desc), device_min_tris, C.byref(h)))
        else:
            check("psy_scene_create", lib().psy_scene_create(C.byref(desc), C.byref(h)))
        self.handle = h
        self._keep = None
        self.mesh_count = len(asm.meshes)
        self.tri_counts = [m.indices.shape[0] for m in asm.meshes]
        self.instance_count = len(asm.instances)

    def info(self) -> dict:
        i = _lib.SceneInfo()
        check("psy_scene_get_info", lib().psy_scene_get_info(self.handle, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _lib.SceneInfo._fields_}

    def dump_bvh(self, mesh: int) -> tuple[np.ndarray, np.ndarray]:
        """Canonical dump of a mesh BVH4 (mesh >= 0) or the instance BVH
        (mesh = -1), and the object order of its leaves."""
        n = lib().psy_scene_dump_bvh(self.handle, mesh, None, 0, None, 0)
        if n < 0:
            check("psy_scene_dump_bvh", int(n))
        out = np.zeros(n, np.uint32)
        order = np.zeros(self.instance_count if mesh < 0 else self.tri_counts[mesh], np.uint32)
        r = lib().psy_scene_dump_bvh(self.handle, mesh, out.ctypes.data_as(_lib._u32p), n,
                                     order.ctypes.data_as(_lib._u32p), len(order))
        if r < 0:
            check("psy_scene_dump_bvh", int(r))
        return out, order

    def close(self):
        if getattr(self, "handle", None):
            lib().psy_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RayBatch:
    """Rays as SoA numpy arrays (RayBatch, ray.rs:24-195)."""

    def __init__(self, ox, oy, oz, dx, dy, dz, time=None, max_t=None, occlusion=None):
        n = len(ox)
        self.ox, self.oy, self.oz = _f32(ox), _f32(oy), _f32(oz)
        self.dx, self.dy, self.dz = _f32(dx), _f32(dy), _f32(dz)
        self.time = _f32(np.zeros(n) if time is None else time)
        self.max_t = _f32(np.full(n, np.inf) if max_t is None else max_t)
        self.flags = None if occlusion is None else np.ascontiguousarray(occlusion, dtype=np.uint8)
        for f in RAY_FIELDS:
            if len(getattr(self, f)) != n:
                raise ValueError(f"ray field {f} has the wrong length")

    def __len__(self):
        return len(self.ox)

    def slice(self, a: int, b: int) -> "RayBatch":
        return RayBatch(*(getattr(self, f)[a:b] for f in RAY_FIELDS),
                        occlusion=None if self.flags is None else self.flags[a:b])

    def soa(self) -> _lib.RaySoA:
        r = _lib.RaySoA()
        for f in RAY_FIELDS:
            setattr(r, f, getattr(self, f).ctypes.data)
        r.flags = None if self.flags is None else self.flags.ctypes.data
        return r


class Hits:
    """Hit records as SoA numpy arrays."""

    def __init__(self, n: int):
        self.kind = np.zeros(n, np.uint8)
        for f in HIT_FLOATS:
            setattr(self, f, np.zeros(n, np.float32))
        for f in HIT_U32:
            setattr(self, f, np.full(n, _lib.PSY_NONE, np.uint32))

    def soa(self) -> _lib.HitSoA:
        h = _lib.HitSoA()
        h.kind = self.kind.ctypes.data
        for f in HIT_FLOATS + HIT_U32:
            setattr(h, f, getattr(self, f).ctypes.data)
        return h

    @classmethod
    def from_records(cls, rec: np.ndarray) -> "Hits":
        """The SoA view of an array of HIT_RECORD."""
        out = cls.__new__(cls)
        out.kind = rec["kind"].astype(np.uint8)
        for f in HIT_FLOATS + HIT_U32:
            setattr(out, f, np.ascontiguousarray(rec[f]))
        return out


class DeviceBuffer:
    """A raw HBM allocation (psy_device_alloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check("psy_device_alloc", lib().psy_device_alloc(max(16, self.nbytes), C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray, offset: int = 0):
        arr = np.ascontiguousarray(arr)
        check("psy_memcpy_h2d", lib().psy_memcpy_h2d(self.ptr + offset, arr.ctypes.data, arr.nbytes))

    def download(self, arr: np.ndarray, offset: int = 0):
        check("psy_memcpy_d2h", lib().psy_memcpy_d2h(arr.ctypes.data, self.ptr + offset, arr.nbytes))

    def free(self):
        if self.ptr:
            lib().psy_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceRays:
    """A RayBatch resident in HBM."""

    def __init__(self, rays: RayBatch):
        n = len(rays)
        self.n = n
        self.buf = DeviceBuffer(n * 4 * 8 + n + 64)
        self.has_flags = rays.flags is not None
        for k, f in enumerate(RAY_FIELDS):
            self.buf.upload(getattr(rays, f), k * n * 4)
        if self.has_flags:
            self.buf.upload(rays.flags, 8 * n * 4)

    def soa(self) -> _lib.RaySoA:
        r = _lib.RaySoA()
        for k, f in enumerate(RAY_FIELDS):
            setattr(r, f, self.buf.ptr + k * self.n * 4)
        r.flags = self.buf.ptr + 8 * self.n * 4 if self.has_flags else None
        return r


class DeviceHits:
    """Hit records resident in HBM."""

    def __init__(self, n: int):
        self.n = n
        self.buf = DeviceBuffer(n * (1 + 4 * 4 + 2 * 4) + 64)

    def soa(self) -> _lib.HitSoA:
        h = _lib.HitSoA()
        base = self.buf.ptr
        n = self.n
        for k, f in enumerate(HIT_FLOATS):
            setattr(h, f, base + k * n * 4)
        for k, f in enumerate(HIT_U32):
            setattr(h, f, base + (4 + k) * n * 4)
        h.kind = base + 6 * n * 4
        return h

    def to_host(self) -> Hits:
        out = Hits(self.n)
        n = self.n
        for k, f in enumerate(HIT_FLOATS + HIT_U32):
            self.buf.download(getattr(out, f), k * n * 4)
        self.buf.download(out.kind, 6 * n * 4)
        return out


class DeviceHitRecords:
    """Hit records (psy_hit_record, 32 bytes per ray) resident in HBM."""

    def __init__(self, n: int):
        self.n = n
        self.buf = DeviceBuffer(n * HIT_RECORD.itemsize + 64)

    def soa(self) -> _lib.HitSoA:
        h = _lib.HitSoA()
        h.records = self.buf.ptr
        return h

    def records(self) -> np.ndarray:
        out = np.zeros(self.n, HIT_RECORD)
        self.buf.download(out)
        return out

    def to_host(self) -> Hits:
        return Hits.from_records(self.records())


class Tracer:
    """Tracer::from_assembly / trace / rays_traced (tracer.rs:15-42)."""

    def __init__(self, scene: Scene, max_batch: int = 1 << 22):
        self.scene = scene
        h = C.c_void_p()
        check("psy_tracer_create", lib().psy_tracer_create(scene.handle, max_batch, C.byref(h)))
        self.handle = h

    @classmethod
    def from_assembly(cls, scene: Scene, max_batch: int = 1 << 22) -> "Tracer":
        return cls(scene, max_batch)

    def trace(self, rays: RayBatch, records: bool = False) -> Hits:
        """Traces host rays; `records` has the library write 32-byte hit
        records (returned as their SoA view) instead of arrays."""
        if records:
            rec = np.zeros(len(rays), HIT_RECORD)
            hs = _lib.HitSoA()
            hs.records = rec.ctypes.data
            rs = rays.soa()
            check("psy_trace", lib().psy_trace(self.handle, C.byref(rs), len(rays), C.byref(hs),
                                               _lib.PSY_MEM_HOST))
            return Hits.from_records(rec)
        hits = Hits(len(rays))
        rs, hs = rays.soa(), hits.soa()
        check("psy_trace", lib().psy_trace(self.handle, C.byref(rs), len(rays), C.byref(hs), _lib.PSY_MEM_HOST))
        return hits

    def trace_device(self, rays: DeviceRays, hits: DeviceHits, sync: bool = True) -> None:
        rs, hs = rays.soa(), hits.soa()
        if sync:
            check("psy_trace", lib().psy_trace(self.handle, C.byref(rs), rays.n, C.byref(hs), _lib.PSY_MEM_DEVICE))
        else:
            check("psy_trace_async", lib().psy_trace_async(self.handle, C.byref(rs), rays.n, C.byref(hs)))

    def finalize(self, rays: DeviceRays, hits) -> np.ndarray:
        """psy_finalize_hits: the hit geometry (HIT_GEOMETRY rows) of device
        rays and their device hits (arrays or records)."""
        out = DeviceBuffer(rays.n * HIT_GEOMETRY.itemsize + 64)
        rs, hs = rays.soa(), hits.soa()
        check("psy_finalize_hits", lib().psy_finalize_hits(self.handle, C.byref(rs), rays.n, C.byref(hs), out.ptr))
        g = np.zeros(rays.n, HIT_GEOMETRY)
        out.download(g)
        return g

    def trace_counted(self, rays: DeviceRays, hits: DeviceHits) -> dict:
        c = _lib.TraceCounters()
        rs, hs = rays.soa(), hits.soa()
        check("psy_trace_counted", lib().psy_trace_counted(self.handle, C.byref(rs), rays.n, C.byref(hs), C.byref(c)))
        return {f: getattr(c, f) for f, _ in _lib.TraceCounters._fields_}

    def tuning(self) -> dict:
        t = _lib.TraceTuning()
        check("psy_tracer_get_tuning", lib().psy_tracer_get_tuning(self.handle, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.TraceTuning._fields_}

    def set_tuning(self, kernel=None, waves_per_cu=None, refill_min=None, stack_lds=None,
                   simd_waves=None, leaf_min=None) -> None:
        cur = self.tuning()
        for k, v in dict(kernel=kernel, waves_per_cu=waves_per_cu, refill_min=refill_min,
                         stack_lds=stack_lds, simd_waves=simd_waves, leaf_min=leaf_min).items():
            if v is not None:
                cur[k] = v
        t = _lib.TraceTuning(**cur)
        check("psy_tracer_set_tuning", lib().psy_tracer_set_tuning(self.handle, C.byref(t)))

    def autotune(self, rays: DeviceRays, hits: DeviceHits) -> dict:
        """Times the candidate launch shapes on `rays`, keeps the fastest."""
        t = _lib.TraceTuning()
        rs, hs = rays.soa(), hits.soa()
        check("psy_tracer_autotune", lib().psy_tracer_autotune(self.handle, C.byref(rs), rays.n, C.byref(hs),
                                                               C.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.TraceTuning._fields_}

    def sync(self) -> None:
        check("psy_tracer_sync", lib().psy_tracer_sync(self.handle))

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check("psy_tracer_last_kernel_ms", lib().psy_tracer_last_kernel_ms(self.handle, C.byref(ms)))
        return ms.value

    def rays_traced(self) -> int:
        return int(lib().psy_tracer_rays_traced(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().psy_tracer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
