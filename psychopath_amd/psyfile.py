""".psy scene files over the C ABI (include/psy_file.h).

    read(text)    ~ DataTree::from_str + parse_scene for every Scene node
                    (src/parse/data_tree.rs:22-45, src/parse/psy.rs:99-200,
                    src/main.rs:235-251); returns RenderScene objects that
                    psychopath_amd.render.build() turns into renderers
    tree(text)    the data tree in the canonical line form of
                    psy_file_dump_tree
    write(scene)  a RenderScene as .psy text in the layout the reader (and
                    the reference's parser) accepts: shaders, objects and
                    instances inside the root Assembly

The reader is C++ in the library; this module only copies its descriptors
into numpy arrays.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, lib
from .scenes import COLOR_BLACKBODY, COLOR_TEMPERATURE, COLOR_XYZ, RenderScene, SubAssembly


class PsyFileError(ValueError):
    """A malformed scene file: `kind` and `offset` as the reference's
    PsyParseError / DataTree ParseError report them."""

    def __init__(self, message: str):
        super().__init__(message)
        head = message.split(" at byte ", 1)
        self.kind = head[0]
        try:
            self.offset = int(head[1].split(" ", 1)[0])
        except (IndexError, ValueError):
            self.offset = -1


def _text_bytes(text) -> bytes:
    return text.encode("utf-8") if isinstance(text, str) else bytes(text)


def tree(text) -> str:
    b = _text_bytes(text)
    n = lib().psy_file_dump_tree(b, len(b), None, 0)
    if n < 0:
        raise PsyFileError(lib().psy_last_error().decode(errors="replace"))
    buf = C.create_string_buffer(int(n) + 1)
    m = lib().psy_file_dump_tree(b, len(b), buf, int(n))
    if m != n:
        raise RuntimeError("psy_file_dump_tree changed length")
    return buf.raw[:n].decode("utf-8")


def _color(c: _lib.Color) -> tuple:
    return (int(c.kind), float(c.a), float(c.b), float(c.c))


def _colours(n, ptr, one) -> tuple | list:
    """One colour as a tuple, time samples as a list of tuples."""
    return [_color(ptr[i]) for i in range(n)] if n else _color(one)


def _array(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def read(text) -> list[RenderScene]:
    """Every Scene of a .psy text as a RenderScene (settings, camera, world,
    shaders, meshes, lights and instances), plus `name` and `output_path`
    attributes.  Raises PsyFileError on a malformed file."""
    b = _text_bytes(text)
    h = C.c_void_p()
    rc = lib().psy_file_parse(b, len(b), C.byref(h))
    if rc == _lib.PSY_E_PARSE:
        raise PsyFileError(lib().psy_last_error().decode(errors="replace"))
    check("psy_file_parse", rc)
    try:
        out = []
        for i in range(lib().psy_file_scene_count(h)):
            fs = _lib.FileScene()
            check("psy_file_get_scene", lib().psy_file_get_scene(h, i, C.byref(fs)))
            out.append(_scene_from(fs))
        return out
    finally:
        lib().psy_file_destroy(h)


def _assembly_from(d, asm) -> None:
    """The meshes, lights, instances and (nested, as in the file)
    sub-assemblies of the assembly descriptor `d` into `asm` (a RenderScene
    or a SubAssembly)."""
    for i in range(d.mesh_count):
        m = d.meshes[i]
        v = _array(m.verts, 3 * m.vert_count * m.time_samples, np.float32).reshape(m.time_samples, m.vert_count, 3)
        t = _array(m.indices, 3 * m.tri_count, np.uint32).reshape(m.tri_count, 3)
        if m.normals:
            nrm = _array(m.normals, 3 * m.vert_count * m.time_samples, np.float32).reshape(v.shape)
            asm.meshes.append((v, t, nrm))
        else:
            asm.meshes.append((v, t))
    for i in range(d.rect_light_count):
        r = d.rect_lights[i]
        if r.dims_count:
            wh = _array(r.dims, 2 * r.dims_count, np.float32).reshape(-1, 2)
            w, h = [float(x) for x in wh[:, 0]], [float(x) for x in wh[:, 1]]
        else:
            w, h = float(r.width), float(r.height)
        asm.rect_lights.append((w, h, _colours(r.color_count, r.colors, r.color)))
    for i in range(d.sphere_light_count):
        s = d.sphere_lights[i]
        rad = [float(x) for x in _array(s.radii, s.radius_count, np.float32)] if s.radius_count else float(s.radius)
        asm.sphere_lights.append((rad, _colours(s.color_count, s.colors, s.color)))
    xf = _array(d.xforms, 12 * d.xform_count, np.float32).reshape(d.xform_count, 12)
    kinds = {_lib.PSY_OBJECT_MESH: "mesh", _lib.PSY_OBJECT_RECT_LIGHT: "rect", _lib.PSY_OBJECT_SPHERE_LIGHT: "sphere",
             _lib.PSY_OBJECT_ASSEMBLY: "assembly"}
    for i in range(d.instance_count):
        inst = d.instances[i]
        asm.instances.append(dict(
            kind=kinds[inst.kind], index=int(inst.mesh),
            shader=None if inst.shader == _lib.PSY_NONE else int(inst.shader),
            xforms=xf[inst.xform_start:inst.xform_start + inst.xform_count].copy() if inst.xform_count else None))
    asm.sub_assemblies = []
    for i in range(d.assembly_count):
        sub = SubAssembly()
        _assembly_from(d.assemblies[i], sub)
        asm.sub_assemblies.append(sub)


def _scene_from(fs: _lib.FileScene) -> RenderScene:
    sc = RenderScene()
    sc.name = fs.name.decode("utf-8")
    sc.output_path = fs.output_path.decode("utf-8")
    _assembly_from(fs.scene, sc)
    w = fs.world
    sc.shaders = [(int(w.shaders[i].kind), _color(w.shaders[i].color), float(w.shaders[i].roughness),
                   float(w.shaders[i].fresnel)) for i in range(w.shader_count)]
    sc.background = _color(w.background)
    sc.distant_lights = []
    for i in range(w.distant_light_count):
        dl = w.distant_lights[i]
        rad = [float(x) for x in _array(dl.radii, dl.radius_count, np.float32)] if dl.radius_count else float(dl.radius)
        if dl.direction_count:
            dirs = [tuple(float(x) for x in v) for v in _array(dl.directions, 3 * dl.direction_count,
                                                                np.float32).reshape(-1, 3)]
        else:
            dirs = tuple(float(x) for x in dl.direction)
        sc.distant_lights.append((rad, dirs, _colours(dl.color_count, dl.colors, dl.color)))
    cam = fs.camera

    def lens(n, p, scalar):
        # One sample as a number, time samples as a list.
        return float(scalar) if n <= 1 else [float(v) for v in _array(p, n, np.float32)]

    sc.camera = dict(xforms=_array(cam.xforms, 12 * cam.xform_count, np.float32).reshape(cam.xform_count, 12),
                     fov=lens(cam.fov_count, cam.fovs, cam.fov),
                     aperture=lens(cam.aperture_count, cam.aperture_radii, cam.aperture_radius),
                     focus=lens(cam.focus_count, cam.focus_distances, cam.focus_distance))
    st = fs.settings
    sc.settings = dict(width=int(st.width), height=int(st.height), spp=int(st.spp), max_bounces=int(st.max_bounces),
                       seed=int(st.seed))
    return sc


# ------------------------------------------------------------------ writer

def _g(x) -> str:
    """Shortest text that reads back as the same f32 (9 significant digits
    always suffice)."""
    x = float(np.float32(x))
    for p in range(6, 10):
        s = "%.*g" % (p, x)
        if float(np.float32(float(s))) == x:
            return s
    return "%.9g" % x


def _matrix(xf) -> str:
    """12 values a..l as the 16 column-major values parse_matrix reads
    (src/parse/psy.rs:556-572)."""
    a, b, c, d, e, f, g, h, i, j, k, l = (float(v) for v in xf)
    return " ".join(_g(v) for v in (a, e, i, 0.0, b, f, j, 0.0, c, g, k, 0.0, d, h, l, 1.0))


def _xyz_to_rec709_e(x, y, z):
    from .scenes import rec709_e_to_xyz_matrix
    return np.linalg.solve(rec709_e_to_xyz_matrix(), np.array([x, y, z], np.float64))


def _color_text(c) -> str:
    kind, a, b, cc = c
    if kind == COLOR_XYZ:
        r, g, bb = _xyz_to_rec709_e(a, b, cc)
        return "rec709, %s %s %s" % (_g(r), _g(g), _g(bb))
    if kind == COLOR_BLACKBODY:
        return "blackbody, %s %s" % (_g(a), _g(b))
    if kind == COLOR_TEMPERATURE:
        return "color_temperature, %s %s" % (_g(a), _g(b))
    raise ValueError(f"unknown colour kind {kind}")


def write(sc: RenderScene, name: str = "Scene", output: str = "render.png") -> str:
    """`sc` as .psy text.  Transforms and vertices read back bit for bit;
    XYZ colours are written as equal-energy Rec.709 and read back within an
    ulp or so.  The bounce limit is not part of the format."""
    from .scenes import SHADER_EMIT, SHADER_GGX, SHADER_LAMBERT
    out = [f"Scene ${name} {{\n"]
    out.append(f'    Output {{\n        Path ["{output}"]\n    }}\n')
    st = sc.settings
    out.append("    RenderSettings {\n"
               f"        Resolution [{st['width']} {st['height']}]\n"
               f"        SamplesPerPixel [{st['spp']}]\n"
               f"        Seed [{st['seed']}]\n    }}\n")
    cam = sc.camera
    out.append("    Camera {\n")
    for v in np.atleast_1d(cam["fov"]):
        out.append(f"        Fov [{_g(math.degrees(float(v)))}]\n")
    for v in np.atleast_1d(cam.get("focus", 1.0)):
        out.append(f"        FocalDistance [{_g(v)}]\n")
    for v in np.atleast_1d(cam.get("aperture", 0.0)):
        out.append(f"        ApertureRadius [{_g(v)}]\n")
    for xf in np.asarray(cam["xforms"], np.float32).reshape(-1, 12):
        out.append(f"        Transform [{_matrix(xf)}]\n")
    out.append("    }\n")
    out.append("    World {\n        BackgroundShader {\n            Type [Color]\n"
               f"            Color [{_color_text(sc.background)}]\n        }}\n")
    for radius, direction, col in sc.distant_lights:
        out.append("        DistantDiskLight {\n")
        out += [f"            Radius [{_g(r)}]\n" for r in np.atleast_1d(radius)]
        dirs = direction if isinstance(direction, list) else [direction]
        out += [f"            Direction [{' '.join(_g(v) for v in d)}]\n" for d in dirs]
        out += [f"            Color [{_color_text(c)}]\n" for c in (col if isinstance(col, list) else [col])]
        out.append("        }\n")
    out.append("    }\n    Assembly {\n")
    for i, (kind, col, rough, fres) in enumerate(sc.shaders):
        names = {SHADER_LAMBERT: "Lambert", SHADER_GGX: "GGX", SHADER_EMIT: "Emit"}
        if kind not in names:
            raise ValueError(f"shader {i}: the .psy format has only Lambert, GGX and Emit shaders")
        t = names[kind]
        out.append(f"        SurfaceShader $shader{i} {{\n            Type [{t}]\n"
                   f"            Color [{_color_text(col)}]\n")
        if kind == SHADER_GGX:
            out.append(f"            Roughness [{_g(rough)}]\n            Fresnel [{_g(fres)}]\n")
        out.append("        }\n")
    for i, mesh in enumerate(sc.meshes):
        v = np.asarray(mesh[0], np.float32)
        if v.ndim == 2:
            v = v[None]
        t = np.asarray(mesh[1], np.uint32).reshape(-1, 3)
        out.append(f"        MeshSurface $mesh{i} {{\n")
        for s in range(v.shape[0]):
            out.append("            Vertices [" + " ".join(_g(x) for x in v[s].ravel()) + "]\n")
        if len(mesh) > 2 and mesh[2] is not None:
            nrm = np.asarray(mesh[2], np.float32).reshape(v.shape)
            for s in range(nrm.shape[0]):
                out.append("            Normals [" + " ".join(_g(x) for x in nrm[s].ravel()) + "]\n")
        out.append("            FaceVertCounts [" + " ".join("3" for _ in range(len(t))) + "]\n")
        out.append("            FaceVertIndices [" + " ".join(str(int(x)) for x in t.ravel()) + "]\n        }\n")
    for i, (w, h, col) in enumerate(sc.rect_lights):
        out.append(f"        RectangleLight $rect{i} {{\n")
        ws, hs = np.broadcast_arrays(np.atleast_1d(w), np.atleast_1d(h))
        out += [f"            Dimensions [{_g(a)} {_g(b)}]\n" for a, b in zip(ws, hs)]
        out += [f"            Color [{_color_text(c)}]\n" for c in (col if isinstance(col, list) else [col])]
        out.append("        }\n")
    for i, (r, col) in enumerate(sc.sphere_lights):
        out.append(f"        SphereLight $sphere{i} {{\n")
        out += [f"            Radius [{_g(x)}]\n" for x in np.atleast_1d(r)]
        out += [f"            Color [{_color_text(c)}]\n" for c in (col if isinstance(col, list) else [col])]
        out.append("        }\n")
    for inst in sc.instances:
        data = {"mesh": "mesh", "rect": "rect", "sphere": "sphere"}[inst["kind"]] + str(inst["index"])
        out.append(f"        Instance {{\n            Data [${data}]\n")
        if inst.get("shader") is not None:
            out.append(f"            SurfaceShaderBind [$shader{inst['shader']}]\n")
        xfs = inst.get("xforms")
        if xfs is not None:
            for xf in np.asarray(xfs, np.float32).reshape(-1, 12):
                out.append(f"            Transform [{_matrix(xf)}]\n")
        out.append("        }\n")
    out.append("    }\n}\n")
    return "".join(out)
