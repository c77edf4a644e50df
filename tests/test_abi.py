"""The C ABI library loads, exports every symbol include/psy_rt.h declares,
and rejects bad input with error codes (no GPU needed)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from psychopath_amd import _lib
from psychopath_amd.tracer import Assembly, Mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        with open(os.path.join(ROOT, "include", h)) as f:
            src = f.read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(psy_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # and the binding covers all of them
    assert sorted(_lib.exported_symbols()) == names


def test_version_and_error_string():
    L = _lib.lib()
    assert b"gfx950" in L.psy_version()
    rc = L.psy_scene_create(None, C.byref(C.c_void_p()))
    assert rc == -1
    assert b"NULL" in L.psy_last_error()


def mesh_desc(verts, tris, samples=1):
    md = _lib.MeshDesc()
    md.vert_count = verts.shape[-2]
    md.time_samples = samples
    md.verts = verts.ctypes.data_as(_lib._fp)
    md.normals = None
    md.tri_count = tris.shape[0]
    md.indices = tris.ctypes.data_as(_lib._u32p)
    return md


def create(meshes, insts, xforms=None):
    L = _lib.lib()
    mds = (_lib.MeshDesc * len(meshes))(*meshes)
    ids = (_lib.InstanceDesc * max(1, len(insts)))(*insts)
    xf = np.zeros((1, 12), np.float32) if xforms is None else xforms
    desc = _lib.SceneDesc(len(meshes), mds, len(insts), ids, 0 if xforms is None else len(xforms),
                          xf.ctypes.data_as(_lib._fp))
    h = C.c_void_p()
    rc = L.psy_scene_create(C.byref(desc), C.byref(h))
    if rc == 0:
        L.psy_scene_destroy(h)
    return rc, L.psy_last_error().decode()


def test_rejects_bad_scenes():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    bad = np.array([[0, 1, 3]], np.uint32)
    ok_inst = _lib.InstanceDesc(0, _lib.PSY_NONE, 0, 0)
    assert create([mesh_desc(v, t)], [ok_inst])[0] == 0
    rc, msg = create([mesh_desc(v, bad)], [ok_inst])
    assert rc == -1 and "out of range" in msg
    rc, msg = create([mesh_desc(v, t, samples=0)], [ok_inst])
    assert rc == -2 and "time_samples" in msg
    rc, msg = create([mesh_desc(v, t)], [_lib.InstanceDesc(3, _lib.PSY_NONE, 0, 0)])
    assert rc == -1 and "object out of range" in msg
    rc, msg = create([mesh_desc(v, t)], [_lib.InstanceDesc(0, _lib.PSY_NONE, 0, 2)], np.zeros((1, 12), np.float32))
    assert rc == -1 and "transform range" in msg


def test_scene_info_and_dump_on_host():
    from psychopath_amd import scenes
    v, f = scenes.displaced_sphere(3)
    a = Assembly()
    a.add_mesh(Mesh(v, f))
    a.add_instance(0)
    s = a.build()
    info = s.info()
    assert info["tri_count"] == 256
    assert info["node_bytes"] == 128 and info["tri_bytes"] == 48 and info["motion_bounds_bytes"] == 96
    # nodes, triangle records, mesh and instance records, the instance
    # position map, the triangle -> record map and the 8 x 48 traversal
    # order table
    assert info["device_bytes"] == info["node_count"] * 128 + 256 * 48 + 32 + 32 + 4 + 256 * 4 + 384
    d, order = s.dump_bvh(0)
    assert sorted(order.tolist()) == list(range(256))
    with pytest.raises(_lib.PsyError):
        L = _lib.lib()
        _lib.check("psy_scene_dump_bvh", int(L.psy_scene_dump_bvh(s.handle, 5, None, 0, None, 0)))


def test_null_tracer_calls_fail_cleanly():
    L = _lib.lib()
    assert L.psy_trace(None, None, 0, None, 0) == -1
    assert L.psy_tracer_sync(None) == -1
    assert L.psy_tracer_rays_traced(None) == 0
    assert L.psy_tracer_create(None, 16, C.byref(C.c_void_p())) == -1


def light_scene(rects=(), spheres=(), insts=(), xforms=None):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    L = _lib.lib()
    md = mesh_desc(v, t)
    mds = (_lib.MeshDesc * 1)(md)
    ids = (_lib.InstanceDesc * max(1, len(insts)))(*insts)
    rl = (_lib.RectLightDesc * max(1, len(rects)))(*rects)
    sl = (_lib.SphereLightDesc * max(1, len(spheres)))(*spheres)
    xf = np.zeros((1, 12), np.float32) if xforms is None else np.ascontiguousarray(xforms, np.float32)
    desc = _lib.SceneDesc(1, mds, len(insts), ids, 0 if xforms is None else len(xf), xf.ctypes.data_as(_lib._fp),
                          len(rects), rl, len(spheres), sl)
    h = C.c_void_p()
    rc = L.psy_scene_create(C.byref(desc), C.byref(h))
    if rc == 0:
        L.psy_scene_destroy(h)
    return rc, L.psy_last_error().decode()


def test_light_objects_validate():
    white = _lib.Color(_lib.PSY_COLOR_XYZ, 1.0, 1.0, 1.0)
    ident = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float32)
    rect = _lib.RectLightDesc(1.0, 2.0, white)
    sph = _lib.SphereLightDesc(0.5, white)
    mesh_inst = _lib.InstanceDesc(0, 0, 0, 0, _lib.PSY_OBJECT_MESH)
    rect_inst = _lib.InstanceDesc(0, _lib.PSY_NONE, 0, 1, _lib.PSY_OBJECT_RECT_LIGHT)
    sph_inst = _lib.InstanceDesc(0, _lib.PSY_NONE, 0, 1, _lib.PSY_OBJECT_SPHERE_LIGHT)
    assert light_scene([rect], [sph], [mesh_inst, rect_inst, sph_inst], ident)[0] == 0
    rc, msg = light_scene([rect], [], [_lib.InstanceDesc(0, _lib.PSY_NONE, 0, 0, _lib.PSY_OBJECT_RECT_LIGHT)])
    assert rc == -1 and "need a transform" in msg
    rc, msg = light_scene([], [sph], [_lib.InstanceDesc(1, _lib.PSY_NONE, 0, 1, _lib.PSY_OBJECT_SPHERE_LIGHT)], ident)
    assert rc == -1 and "object out of range" in msg
    rc, msg = light_scene([_lib.RectLightDesc(0.0, 1.0, white)], [], [rect_inst], ident)
    assert rc == -1 and "bad size" in msg
    rc, msg = light_scene([], [_lib.SphereLightDesc(1.0, _lib.Color(7, 1, 1, 1))], [sph_inst], ident)
    assert rc == -1 and "bad colour" in msg
    rc, msg = light_scene([rect], [], [_lib.InstanceDesc(0, _lib.PSY_NONE, 0, 1, 9)], ident)
    assert rc == -1 and "unknown object kind" in msg
    # Time samples: sizes must be positive, colours of one representation,
    # lists at most 64 long and not NULL.
    radii = np.array([0.5, 0.0], np.float32)
    rc, msg = light_scene([], [_lib.SphereLightDesc(1.0, white, 2, radii.ctypes.data_as(_lib._fp))], [sph_inst], ident)
    assert rc == -1 and "bad radius" in msg
    mixed = (_lib.Color * 2)(white, _lib.Color(_lib.PSY_COLOR_BLACKBODY, 5000.0, 1.0, 0.0))
    rc, msg = light_scene([_lib.RectLightDesc(1.0, 1.0, white, 0, None, 2, mixed)], [], [rect_inst], ident)
    assert rc == -1 and "different representations" in msg
    rc, msg = light_scene([_lib.RectLightDesc(1.0, 1.0, white, 2, None)], [], [rect_inst], ident)
    assert rc == -1 and "NULL" in msg
    many = np.ones(65 * 2, np.float32)
    rc, msg = light_scene([_lib.RectLightDesc(1.0, 1.0, white, 65, many.ctypes.data_as(_lib._fp))], [], [rect_inst],
                          ident)
    assert rc == -2 and "too many" in msg
    dims = np.array([1.0, 2.0, 3.0, 0.5], np.float32)
    ok = _lib.RectLightDesc(9.0, 9.0, white, 2, dims.ctypes.data_as(_lib._fp), 2,
                            (_lib.Color * 2)(white, _lib.Color(_lib.PSY_COLOR_XYZ, 2.0, 2.0, 2.0)))
    assert light_scene([ok], [], [rect_inst], ident)[0] == 0


def test_null_renderer_calls_fail_cleanly():
    L = _lib.lib()
    assert L.psy_renderer_create(None, None, 0, C.byref(C.c_void_p())) == -1
    assert L.psy_render(None, None, None, None, 0, None) == -1
    assert L.psy_renderer_destroy(None) == 0


def test_null_server_calls_fail_cleanly():
    L = _lib.lib()
    assert L.psy_server_create(None, 0, None, C.byref(C.c_void_p())) == -1
    assert L.psy_server_submit(None, 0, None, None, None) == -1
    assert L.psy_server_wait(None, 1, None, None) == -1
    assert L.psy_server_destroy(None) == 0


def test_nested_assembly_descriptions():
    """Sub-assemblies in the scene description: flattened on the host, with
    bad indices, cycles of depth and over-deep light nesting refused."""
    from psychopath_amd import render, scenes
    from psychopath_amd.tracer import Assembly, Mesh
    sc = scenes.nested_lights_scene(8, 8, 1, 1, 0)
    info = render.assembly_of(sc).build().info()
    # 7 room meshes and the lamp's box; every lamp instance path flattened
    assert info["tri_count"] == 46
    v, f = scenes.displaced_sphere(1)
    # an instance naming a missing sub-assembly
    a = Assembly()
    a.add_mesh(Mesh(v, f))
    a.add_instance(0, kind=_lib.PSY_OBJECT_ASSEMBLY)
    with pytest.raises(_lib.PsyError, match="object out of range"):
        a.build()
    # nesting deeper than 64 levels
    leaf = Assembly()
    leaf.add_mesh(Mesh(v, f))
    leaf.add_instance(0)
    cur = leaf
    for _ in range(70):
        up = Assembly()
        up.add_assembly(cur)
        up.add_instance(0, kind=_lib.PSY_OBJECT_ASSEMBLY)
        cur = up
    with pytest.raises(_lib.PsyError, match="deeper than 64"):
        cur.build()
    # lights four sub-assembly levels down
    from psychopath_amd.tracer import SphereLight
    lit = Assembly()
    lit.add_sphere_light(SphereLight(0.1, (0, 1.0, 1.0, 1.0)))
    lit.add_instance(0, np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32), kind=_lib.PSY_OBJECT_SPHERE_LIGHT)
    cur = lit
    for _ in range(4):
        up = Assembly()
        up.add_assembly(cur)
        up.add_instance(0, kind=_lib.PSY_OBJECT_ASSEMBLY)
        cur = up
    with pytest.raises(_lib.PsyError, match="lights nested"):
        cur.build()
    # three levels are accepted
    cur = lit
    for _ in range(3):
        up = Assembly()
        up.add_assembly(cur)
        up.add_instance(0, kind=_lib.PSY_OBJECT_ASSEMBLY)
        cur = up
    assert cur.build().info()["tri_count"] == 0


def test_noalloc_entry_point_checks_before_touching_the_device():
    from psychopath_amd import scenes
    v, f = scenes.displaced_sphere(2)
    a = Assembly()
    a.add_mesh(Mesh(v, f))
    a.add_instance(0)
    s = a.build()  # never uploaded: no tracer was created
    L = _lib.lib()
    fake = 16  # non-NULL addresses the checks never dereference
    rs = _lib.RaySoA(fake, fake, fake, fake, fake, fake, fake, fake, None)
    hs = _lib.HitSoA(None, None, None, None, None, None, None, 16)
    assert L.psy_trace_noalloc(None, C.byref(rs), 1, C.byref(hs), None, None) == _lib.PSY_E_INVALID
    assert L.psy_trace_noalloc(s.handle, C.byref(rs), _lib.PSY_NOALLOC_MAX_RAYS + 1, C.byref(hs), None,
                               None) == _lib.PSY_E_RANGE
    assert "2^31" in L.psy_last_error().decode()
    assert L.psy_trace_noalloc(s.handle, C.byref(rs), 1, C.byref(hs), None, None) == _lib.PSY_E_STATE
    assert "not resident" in L.psy_last_error().decode()
    bad = _lib.TraceTuning(_lib.PSY_KERNEL_PERSISTENT, 0, 16, 16, 4, 0)
    assert L.psy_trace_noalloc(s.handle, C.byref(rs), 1, C.byref(hs), C.byref(bad), None) == _lib.PSY_E_RANGE
