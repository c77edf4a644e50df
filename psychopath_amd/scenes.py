"""Seeded, deterministic synthetic scenes and ray sets.

Ray-cast workload (1M triangles, 16M rays):
  * mesh: a regular tetrahedron subdivided `levels` times (4 * 4**levels
    triangles; 9 levels give 1,048,576), vertices projected to the unit
    sphere at every level, then displaced along the sphere normal by
    `amplitude * perlin3(p)` (Perlin's 2002 improved noise, permutation
    shuffled with seed 1);
  * incoherent rays: origins uniform in the mesh bounds (seed 2), directions
    `uniform_sample_sphere(u, v)` (src/sampling/monte_carlo.rs:49-56,
    seed 3), time uniform in [0,1) (seed 4), max_t = +inf;
  * shadow rays: origin and end point uniform in the bounds (seeds 5, 6),
    dir = end - origin, max_t = 1.0, occlusion flag set (the shape of
    src/renderer.rs:568-574);
  * coherent rays: a W x H pinhole camera at pixel centres, built as
    Camera::generate_ray does (src/camera.rs:74-102) with lens (0.5, 0.5).

Random numbers come from SplitMix64: value i of stream `seed` is the i-th
output of a SplitMix64 generator seeded with `seed`; a float in [0,1) is its
top 24 bits times 2**-24.
"""

from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+n-1 of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, start: int = 0) -> np.ndarray:
    """n floats in [0,1) with 24 random bits each (exact in float32)."""
    return ((splitmix64(seed, n, start) >> np.uint64(40)).astype(np.float32)
            * np.float32(1.0 / 16777216.0)).astype(np.float32)


# ---------------------------------------------------------------- geometry

def tetra_sphere(levels: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit-sphere mesh from a subdivided regular tetrahedron (float64 verts,
    uint32 triangles, outward counter-clockwise winding)."""
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) / np.sqrt(3.0)
    f = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], np.int64)
    # orient every face outward
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = np.einsum("ij,ij->i", n, v[f].mean(axis=1)) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    for _ in range(levels):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(3, -1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = inv[0] + base, inv[1] + base, inv[2] + base
        f = np.concatenate([
            np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return v, f.astype(np.uint32)


def perlin_permutation(seed: int) -> np.ndarray:
    """0..255 shuffled by Fisher-Yates driven by SplitMix64(seed), doubled."""
    p = np.arange(256, dtype=np.int64)
    r = splitmix64(seed, 255)
    for k, i in enumerate(range(255, 0, -1)):
        j = int(r[k] % np.uint64(i + 1))
        p[i], p[j] = p[j], p[i]
    return np.concatenate([p, p])


def perlin3(pts: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """Perlin's improved noise (2002) at points [N,3] (float64)."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    fx, fy, fz = np.floor(x), np.floor(y), np.floor(z)
    X = fx.astype(np.int64) & 255
    Y = fy.astype(np.int64) & 255
    Z = fz.astype(np.int64) & 255
    x, y, z = x - fx, y - fy, z - fz

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def grad(h, x, y, z):
        h = h & 15
        u = np.where(h < 8, x, y)
        v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -u, u) + np.where(h & 2, -v, v)

    def lerp(t, a, b):
        return a + t * (b - a)

    u, v, w = fade(x), fade(y), fade(z)
    p = perm
    A = p[X] + Y
    AA, AB = p[A] + Z, p[A + 1] + Z
    B = p[X + 1] + Y
    BA, BB = p[B] + Z, p[B + 1] + Z
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))))


def displaced_sphere(levels: int = 9, amplitude: float = 0.1, seed: int = 1,
                     frequency: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """The ray-cast mesh: float32 verts [V,3] and uint32 triangles [T,3]."""
    v, f = tetra_sphere(levels)
    perm = perlin_permutation(seed)
    d = perlin3(v * frequency, perm)
    v = v * (1.0 + amplitude * d)[:, None]
    v = v.astype(np.float32) + np.float32(0.0)  # no negative zeros
    return v, f


# -------------------------------------------------------------------- rays

def _sphere_dirs(u: np.ndarray, v: np.ndarray):
    """uniform_sample_sphere (monte_carlo.rs:49-56) in float32."""
    z = np.float32(1.0) - np.float32(2.0) * u
    r = np.sqrt(np.maximum(np.float32(1.0) - z * z, np.float32(0.0)))
    phi = np.float32(2.0 * np.pi) * v
    return (r * np.cos(phi)).astype(np.float32), (r * np.sin(phi)).astype(np.float32), z.astype(np.float32)


def incoherent_rays(lo, hi, n: int, seeds=(2, 3, 4)) -> dict:
    lo = np.asarray(lo, np.float32)
    ext = (np.asarray(hi, np.float32) - lo).astype(np.float32)
    uo = uniform01(seeds[0], 3 * n).reshape(n, 3)
    o = (lo + ext * uo).astype(np.float32)
    ud = uniform01(seeds[1], 2 * n).reshape(n, 2)
    dx, dy, dz = _sphere_dirs(ud[:, 0], ud[:, 1])
    t = uniform01(seeds[2], n)
    return dict(ox=o[:, 0].copy(), oy=o[:, 1].copy(), oz=o[:, 2].copy(), dx=dx, dy=dy, dz=dz, time=t,
                max_t=np.full(n, np.inf, np.float32))


def shadow_rays(lo, hi, n: int, seeds=(5, 6)) -> dict:
    lo = np.asarray(lo, np.float32)
    ext = (np.asarray(hi, np.float32) - lo).astype(np.float32)
    a = (lo + ext * uniform01(seeds[0], 3 * n).reshape(n, 3)).astype(np.float32)
    b = (lo + ext * uniform01(seeds[1], 3 * n).reshape(n, 3)).astype(np.float32)
    d = (b - a).astype(np.float32)
    return dict(ox=a[:, 0].copy(), oy=a[:, 1].copy(), oz=a[:, 2].copy(), dx=d[:, 0].copy(), dy=d[:, 1].copy(),
                dz=d[:, 2].copy(), time=np.zeros(n, np.float32), max_t=np.ones(n, np.float32),
                occlusion=np.ones(n, np.uint8))


def coherent_rays(lo, hi, width: int, height: int, fov_deg: float = 40.0) -> dict:
    """Pinhole primary rays at pixel centres looking down +z at the bounds.

    Camera space follows src/camera.rs:74-102 with lens (0.5, 0.5), so the
    origin is the camera position and dir = normalize(x*tfov, y*tfov, 1); the
    camera transform is a translation placing the camera in front of the
    bounds.  Image-plane x spans [-1, 1] across the width and y keeps the
    aspect ratio (src/renderer.rs:248-255 without the pixel filter).
    """
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    c = (lo + hi) / 2
    radius = float(np.linalg.norm(hi - lo) / 2)
    half = np.radians(fov_deg) / 2
    tfov = np.float32(np.sin(half) / np.cos(half))
    dist = radius / np.sin(half)
    cam = np.array([c[0], c[1], c[2] - dist], np.float32)
    aspect = height / width
    px = ((np.arange(width, dtype=np.float32) + np.float32(0.5)) / np.float32(width)) * np.float32(2) - np.float32(1)
    py = (np.float32(1) - ((np.arange(height, dtype=np.float32) + np.float32(0.5)) / np.float32(height))
          * np.float32(2)) * np.float32(aspect)
    X, Y = np.meshgrid(px, py)
    x = (X.ravel() * tfov).astype(np.float32)
    y = (Y.ravel() * tfov).astype(np.float32)
    z = np.ones_like(x)
    inv_len = (np.float32(1.0) / np.sqrt((x * x + y * y) + z * z)).astype(np.float32)
    n = width * height
    return dict(ox=np.full(n, cam[0], np.float32), oy=np.full(n, cam[1], np.float32),
                oz=np.full(n, cam[2], np.float32), dx=(x * inv_len).astype(np.float32),
                dy=(y * inv_len).astype(np.float32), dz=(z * inv_len).astype(np.float32),
                time=np.zeros(n, np.float32), max_t=np.full(n, np.inf, np.float32))


def raycast_workload(levels: int = 9, n_rays: int = 1 << 24):
    """The 1M-triangle ray-cast mesh and its incoherent ray set."""
    verts, tris = displaced_sphere(levels)
    lo, hi = verts.min(axis=0), verts.max(axis=0)
    return verts, tris, incoherent_rays(lo, hi, n_rays)


# ------------------------------------------------------------- render scenes
# Colours are (kind, a, b, c): kind 0 = XYZ (x, y, z), 1 = blackbody
# (temperature K, factor), 2 = normalised temperature (temperature K, factor)
# (Color, src/color.rs:46-59).  Shaders are (kind, colour, roughness, fresnel)
# with kind 0 = Lambert, 1 = GGX, 2 = Emit (SimpleSurfaceShader,
# src/shading/mod.rs:27-40).

COLOR_XYZ, COLOR_BLACKBODY, COLOR_TEMPERATURE = 0, 1, 2
SHADER_LAMBERT, SHADER_GGX, SHADER_EMIT, SHADER_DIELECTRIC = 0, 1, 2, 3


def rec709_e_to_xyz_matrix() -> np.ndarray:
    """RGB (Rec.709 primaries, equal-energy white) to CIE XYZ, derived from
    the primaries' chromaticities: columns are the primaries' XYZ scaled so
    RGB (1,1,1) maps to the white point's XYZ (1,1,1)."""
    prim = np.array([[0.64, 0.33], [0.30, 0.60], [0.15, 0.06]])
    P = np.stack([prim[:, 0] / prim[:, 1], np.ones(3), (1 - prim[:, 0] - prim[:, 1]) / prim[:, 1]])
    S = np.linalg.solve(P, np.ones(3))
    return P * S[None, :]


def rec709(r, g, b):
    x, y, z = rec709_e_to_xyz_matrix() @ np.array([r, g, b], np.float64)
    return (COLOR_XYZ, float(np.float32(x)), float(np.float32(y)), float(np.float32(z)))


def translate_xf(t):
    """World-to-object transform of an object placed at `t` (12 values a..l)."""
    return np.array([1, 0, 0, -t[0], 0, 1, 0, -t[1], 0, 0, 1, -t[2]], np.float32)



class RenderScene:
    """Everything a render needs, in plain arrays: meshes, light objects,
    instances, shaders, world, camera and settings.  Transforms are the 12
    values a..l of Transform::new_from_values; instance transforms map world
    to object space, camera transforms camera to world space."""

    def __init__(self):
        self.meshes = []          # (verts [S,V,3] or [V,3], tris [T,3])
        self.rect_lights = []     # (width, height, colour)
        self.sphere_lights = []   # (radius, colour)
        self.instances = []       # dict(kind="mesh"|"rect"|"sphere", index, shader, xforms)
        self.shaders = []         # (kind, colour, roughness, fresnel; the IOR for a dielectric)
        self.background = (COLOR_XYZ, 0.0, 0.0, 0.0)
        self.distant_lights = []  # (radius, direction (3), colour)
        self.camera = dict(xforms=np.zeros((1, 12), np.float32), fov=1.0, aperture=0.0, focus=1.0)
        self.settings = dict(width=64, height=64, spp=1, max_bounces=2, seed=0)

    def add_mesh_instance(self, verts, tris, shader, xforms=None):
        self.meshes.append((np.asarray(verts, np.float32), np.asarray(tris, np.uint32)))
        self.instances.append(dict(kind="mesh", index=len(self.meshes) - 1, shader=shader, xforms=xforms))


def _quad(p0, p1, p2, p3):
    v = np.array([p0, p1, p2, p3], np.float32)
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _box(center, size, angle_deg):
    """An axis-aligned box of `size` rotated about the vertical (z) axis and
    placed with its base at z = 0: 8 vertices, 12 outward triangles."""
    sx, sy, sz = size
    c = np.cos(np.radians(angle_deg))
    s = np.sin(np.radians(angle_deg))
    corners = []
    for dz in (0.0, sz):
        for dx, dy in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
            x, y = dx * sx, dy * sy
            corners.append((center[0] + c * x - s * y, center[1] + s * x + c * y, dz))
    v = np.array(corners, np.float32) + np.float32(0.0)
    f = np.array([[0, 2, 1], [0, 3, 2],            # bottom
                  [4, 5, 6], [4, 6, 7],            # top
                  [0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5],
                  [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.uint32)
    return v, f


def cornell_box(width=512, height=512, spp=16, max_bounces=4, seed=0) -> RenderScene:
    """The Cornell-box workload: five axis-aligned walls (two triangles each,
    Lambertian white/red/green), a tall and a short box (twelve triangles
    each, rotated 15 and -18 degrees about the vertical), a 0.25 x 0.25
    rectangle light just under the ceiling, and a pinhole camera.  Spectra
    (Rec.709 reflectances and light colour), camera transform, field of view
    and focal distance are those of the reference's example scene
    (example_scenes/cornell_box.psy:10-34,38); its walls and light are
    simplified as the workload describes."""
    sc = RenderScene()
    white = rec709(0.7295, 0.7355, 0.729)
    red = rec709(0.611, 0.0555, 0.062)
    green = rec709(0.117, 0.4125, 0.115)
    sc.shaders = [(SHADER_LAMBERT, white, 0.0, 0.0), (SHADER_LAMBERT, red, 0.0, 0.0),
                  (SHADER_LAMBERT, green, 0.0, 0.0)]
    X0, X1, Y1, Z1 = -5.56, 0.0, 5.592, 5.488
    walls = [
        (_quad((X0, 0, 0), (X1, 0, 0), (X1, Y1, 0), (X0, Y1, 0)), 0),          # floor
        (_quad((X0, 0, Z1), (X0, Y1, Z1), (X1, Y1, Z1), (X1, 0, Z1)), 0),      # ceiling
        (_quad((X0, Y1, 0), (X1, Y1, 0), (X1, Y1, Z1), (X0, Y1, Z1)), 0),      # back
        (_quad((X0, 0, 0), (X0, Y1, 0), (X0, Y1, Z1), (X0, 0, Z1)), 1),        # left, red
        (_quad((X1, 0, 0), (X1, 0, Z1), (X1, Y1, Z1), (X1, Y1, 0)), 2),        # right, green
    ]
    for (v, f), shader in walls:
        sc.add_mesh_instance(v, f, shader)
    v, f = _box((-3.68, 3.51, 0.0), (1.65, 1.65, 3.30), 15.0)
    sc.add_mesh_instance(v, f, 0)
    v, f = _box((-1.86, 1.69, 0.0), (1.65, 1.65, 1.65), -18.0)
    sc.add_mesh_instance(v, f, 0)
    sc.rect_lights.append((0.25, 0.25, rec709(84.300003, 53.800003, 18.5)))
    # The example's light placement (cornell_box.psy:84-86: world to object
    # is a pure translation), lowered 1 cm under the closed ceiling.
    light_xf = np.array([1, 0, 0, 2.779475, 0, 1, 0, -2.794788, 0, 0, 1, -5.478], np.float32)
    sc.instances.append(dict(kind="rect", index=0, shader=None, xforms=light_xf[None]))
    # Camera (cornell_box.psy:14): camera x -> world x, camera y -> world z,
    # camera z (view direction) -> world y.
    sc.camera = dict(xforms=np.array([[1, 0, 0, -2.779998, 0, 0, 1, -8.0, 0, 1, 0, 2.73001]], np.float32),
                     fov=float(np.float32(39.449188 * np.pi / 180.0)), aperture=0.0, focus=10.62)
    sc.settings = dict(width=width, height=height, spp=spp, max_bounces=max_bounces, seed=seed)
    return sc


class SubAssembly:
    """A sub-assembly of a RenderScene: meshes, light objects, instances and
    its own sub-assemblies, instanced by instances of kind "assembly" (index
    into the parent's `sub_assemblies`).  Shader indices refer to the
    scene's one shader list."""

    def __init__(self):
        self.meshes = []
        self.rect_lights = []
        self.sphere_lights = []
        self.instances = []
        self.sub_assemblies = []


def nested_lights_scene(width=96, height=96, spp=8, max_bounces=3, seed=0) -> RenderScene:
    """The Cornell box with lamps in nested assemblies: a "lamp" assembly
    (a small box shade, a sphere light and a rectangle light, and a "bulb"
    sub-assembly holding a tiny sphere light, instanced twice inside the
    lamp) placed three times in the room, once under two-sample motion.
    Light selection then recurses through two levels of sub-assembly light
    trees (scene/assembly.rs:46-141)."""
    sc = cornell_box(width, height, spp, max_bounces, seed)
    bulb = SubAssembly()
    bulb.sphere_lights.append((0.04, (COLOR_BLACKBODY, 3200.0, 4.0e-10, 0.0)))
    bulb.instances.append(dict(kind="sphere", index=0, shader=None, xforms=translate_xf((0.0, 0.0, 0.05))[None]))
    lamp = SubAssembly()
    v, f = _box((0.0, 0.0, 0.0), (0.3, 0.3, 0.2), 10.0)
    lamp.meshes.append((v, f))
    lamp.instances.append(dict(kind="mesh", index=0, shader=0, xforms=None))
    lamp.sphere_lights.append((0.07, rec709(12.0, 10.0, 7.0)))
    lamp.instances.append(dict(kind="sphere", index=0, shader=None, xforms=translate_xf((0.0, 0.0, 0.35))[None]))
    lamp.rect_lights.append((0.12, 0.2, rec709(30.0, 28.0, 20.0)))
    lamp.instances.append(dict(kind="rect", index=0, shader=None,
                               xforms=world_to_object(_rotation((1.0, 0.0, 0.0), 2.2), np.array([0.0, -0.2, 0.28]))[None]))
    lamp.sub_assemblies.append(bulb)
    lamp.instances.append(dict(kind="assembly", index=0, shader=None, xforms=translate_xf((0.1, 0.0, 0.2))[None]))
    lamp.instances.append(dict(kind="assembly", index=0, shader=None,
                               xforms=world_to_object(_rotation((0.0, 0.0, 1.0), 0.7), np.array([-0.1, 0.05, 0.2]))[None]))
    sc.sub_assemblies = [lamp]
    places = [(-4.3, 1.2, 0.0), (-1.0, 4.3, 0.0)]
    for p in places:
        sc.instances.append(dict(kind="assembly", index=0, shader=None, xforms=translate_xf(p)[None]))
    moving = np.stack([world_to_object(_rotation((0.0, 0.0, 1.0), a), np.array([-2.7, 2.4, 3.3 + 0.4 * a]))
                       for a in (0.0, 0.6)])
    sc.instances.append(dict(kind="assembly", index=0, shader=None, xforms=moving))
    return sc


def normal01(seed: int, n: int) -> np.ndarray:
    """Seeded standard normal samples (Box-Muller over uniform01)."""
    u = uniform01(seed, 2 * n).astype(np.float64).reshape(2, n)
    u1 = np.maximum(u[0], 1e-12)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1])


def _rotation(axis, angle):
    """Rodrigues rotation matrix about a unit axis."""
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    C = 1.0 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def world_to_object(rot, pos):
    """12 transform values a..l mapping world to object space for an object
    placed by p_world = rot @ p_object + pos."""
    rt = rot.T
    t = -rt @ pos
    return np.array([rt[0, 0], rt[0, 1], rt[0, 2], t[0], rt[1, 0], rt[1, 1], rt[1, 2], t[1],
                     rt[2, 0], rt[2, 1], rt[2, 2], t[2]], np.float32)


def look_at(pos, target, up=(0.0, 0.0, 1.0)):
    """Camera-to-world transform (camera x right, y up, z forward) of a
    camera at `pos` looking at `target`."""
    pos = np.asarray(pos, np.float64)
    fwd = np.asarray(target, np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    return np.array([right[0], upv[0], fwd[0], pos[0], right[1], upv[1], fwd[1], pos[1],
                     right[2], upv[2], fwd[2], pos[2]], np.float32)


def motion_instances(grid: int = 16, levels: int = 7, lights: int = 16, width: int = 1920, height: int = 1080,
                     spp: int = 64, max_bounces: int = 6, seed: int = 0, aperture: float = 0.05,
                     smooth: bool = False) -> RenderScene:
    """The motion-blur and instancing workload: grid x grid instances of a
    displaced sphere (4 * 4^levels triangles) with two deformation time
    samples (vertex offsets N(0, 0.02)), each instance on a seeded jittered
    grid with three time-sampled transforms (rotation up to 30 degrees,
    translation up to 0.5), `lights` blackbody sphere lights of radius
    0.05-0.2 at 2500-9000 K, and a thin-lens camera."""
    sc = RenderScene()
    v, f = displaced_sphere(levels, amplitude=0.1, seed=1)
    off = normal01(seed * 7 + 11, v.size).reshape(v.shape) * 0.02
    v1 = (v.astype(np.float64) + off).astype(np.float32) + np.float32(0.0)
    vs = np.stack([v, v1])
    sc.meshes.append((vs, f, vertex_normals(vs, f)) if smooth else (vs, f))
    sc.shaders = [(SHADER_LAMBERT, rec709(0.8, 0.8, 0.8), 0.0, 0.0),
                  (SHADER_GGX, rec709(1.0, 0.78, 0.34), 0.2, 0.8),
                  (SHADER_LAMBERT, rec709(0.7, 0.2, 0.15), 0.0, 0.0)]
    n = grid * grid
    u = uniform01(seed * 7 + 1, 9 * n).astype(np.float64).reshape(n, 9)
    spacing = 2.5
    for i in range(n):
        gx, gy = i % grid, i // grid
        base = np.array([(gx - (grid - 1) / 2) * spacing + (u[i, 0] - 0.5) * spacing * 0.4,
                         (gy - (grid - 1) / 2) * spacing + (u[i, 1] - 0.5) * spacing * 0.4, 0.0])
        axis = _sphere_point(u[i, 2], u[i, 3])
        ang = np.radians(30.0) * u[i, 4]
        move = _sphere_point(u[i, 5], u[i, 6]) * 0.5 * u[i, 7]
        xfs = [world_to_object(_rotation(axis, ang * k / 2.0), base + move * k / 2.0) for k in range(3)]
        sc.instances.append(dict(kind="mesh", index=0, shader=i % 3, xforms=np.stack(xfs)))
    lu = uniform01(seed * 7 + 2, 5 * lights).astype(np.float64).reshape(lights, 5)
    half = grid * spacing / 2
    for k in range(lights):
        radius = 0.05 + 0.15 * lu[k, 0]
        temp = 2500.0 + 6500.0 * lu[k, 1]
        sc.sphere_lights.append((radius, (COLOR_TEMPERATURE, temp, 40.0, 0.0)))
        pos = np.array([(lu[k, 2] - 0.5) * 2 * half, (lu[k, 3] - 0.5) * 2 * half, 2.0 + 2.0 * lu[k, 4]])
        sc.instances.append(dict(kind="sphere", index=k, shader=None,
                                 xforms=world_to_object(np.eye(3), pos)[None]))
    cam_pos = np.array([0.0, -half * 1.6, half * 0.9])
    focus = float(np.linalg.norm(cam_pos))
    sc.camera = dict(xforms=look_at(cam_pos, (0.0, 0.0, 0.0))[None], fov=float(np.float32(np.radians(50.0))),
                     aperture=aperture, focus=focus)
    sc.settings = dict(width=width, height=height, spp=spp, max_bounces=max_bounces, seed=seed)
    return sc


def vertex_normals(verts: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """Smooth-shading normals: area-weighted face normals summed per vertex
    and normalised (float64, returned as float32), for each time sample of
    `verts` ([V,3] or [S,V,3])."""
    v = np.asarray(verts, np.float64)
    single = v.ndim == 2
    if single:
        v = v[None]
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    out = np.zeros_like(v)
    for s in range(v.shape[0]):
        fn = np.cross(v[s, t[:, 1]] - v[s, t[:, 0]], v[s, t[:, 2]] - v[s, t[:, 0]])
        for k in range(3):
            np.add.at(out[s], t[:, k], fn)
    out /= np.maximum(np.linalg.norm(out, axis=-1, keepdims=True), 1e-30)
    out = out.astype(np.float32) + np.float32(0.0)
    return out[0] if single else out


def _sphere_point(u, v):
    z = 1.0 - 2.0 * u
    r = np.sqrt(max(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    return np.array([r * np.cos(phi), r * np.sin(phi), z])


def value_noise2(x: np.ndarray, y: np.ndarray, seed: int) -> np.ndarray:
    """Seeded 2D value noise in [-1, 1]: hashed lattice values blended with
    a smoothstep."""
    x0 = np.floor(x)
    y0 = np.floor(y)
    fx = x - x0
    fy = y - y0
    ix = x0.astype(np.int64)
    iy = y0.astype(np.int64)
    ox, oy = ix.min(), iy.min()
    w, h = ix.max() - ox + 2, iy.max() - oy + 2
    lattice = uniform01(seed, int(w * h)).astype(np.float64).reshape(h, w) * 2.0 - 1.0
    gx, gy = ix - ox, iy - oy
    sx = fx * fx * (3.0 - 2.0 * fx)
    sy = fy * fy * (3.0 - 2.0 * fy)
    a = lattice[gy, gx]
    b = lattice[gy, gx + 1]
    c = lattice[gy + 1, gx]
    d = lattice[gy + 1, gx + 1]
    return (a * (1 - sx) + b * sx) * (1 - sy) + (c * (1 - sx) + d * sx) * sy


def heightfield(nx: int = 2048, ny: int = 1280, seed: int = 3, spacing: float = 0.01,
                amplitude: float = 1.5) -> tuple[np.ndarray, np.ndarray]:
    """Seeded terrain: (nx+1) x (ny+1) vertices on a grid of `spacing`,
    heights from four octaves of Perlin noise, two triangles per quad
    (2 * nx * ny triangles)."""
    xs = np.arange(nx + 1, dtype=np.float64) * spacing
    ys = np.arange(ny + 1, dtype=np.float64) * spacing
    gx, gy = np.meshgrid(xs, ys)
    h = np.zeros(gx.shape)
    freq, amp = 0.35, 1.0
    for octave in range(4):
        h += amp * value_noise2(gx * freq, gy * freq, seed * 16 + octave)
        freq *= 2.0
        amp *= 0.5
    v = np.stack([gx.ravel(), gy.ravel(), h.ravel() * amplitude], axis=1).astype(np.float32) + np.float32(0.0)
    i = np.arange(nx, dtype=np.uint32)
    j = np.arange(ny, dtype=np.uint32)
    ii, jj = np.meshgrid(i, j)
    a = (jj * (nx + 1) + ii).ravel()
    b = a + 1
    c = a + (nx + 1)
    d = c + 1
    f = np.empty((2 * a.size, 3), np.uint32)
    f[0::2] = np.stack([a, b, d], axis=1)
    f[1::2] = np.stack([a, d, c], axis=1)
    return v, f


def terrain_scene(nx: int = 2048, ny: int = 1280, width: int = 1024, height: int = 1024, spp: int = 16,
                  max_bounces: int = 4, seed: int = 0) -> RenderScene:
    """The serving workload's terrain: a 2 * nx * ny triangle heightfield,
    Lambertian, lit by a distant disk light (the sun) under a sky-coloured
    background, seen from above at a grazing angle."""
    sc = RenderScene()
    v, f = heightfield(nx, ny)
    sc.meshes.append((v, f))
    sc.shaders = [(SHADER_LAMBERT, rec709(0.35, 0.3, 0.22), 0.0, 0.0)]
    sc.instances.append(dict(kind="mesh", index=0, shader=0, xforms=None))
    sc.background = rec709(0.25, 0.35, 0.6)
    sc.distant_lights.append((0.02, (-0.4, 0.3, -0.6), (COLOR_TEMPERATURE, 5800.0, 100.0, 0.0)))
    lo, hi = v.min(axis=0).astype(np.float64), v.max(axis=0).astype(np.float64)
    center = (lo + hi) / 2
    ext = hi - lo
    cam = center + np.array([0.0, -0.75 * ext[1], 0.35 * ext[0]])
    sc.camera = dict(xforms=look_at(cam, center)[None], fov=float(np.float32(np.radians(55.0))), aperture=0.0,
                     focus=1.0)
    sc.settings = dict(width=width, height=height, spp=spp, max_bounces=max_bounces, seed=seed)
    return sc


def sphere_scene(levels: int = 9, width: int = 512, height: int = 512, spp: int = 16, max_bounces: int = 4,
                 seed: int = 0, smooth: bool = False) -> RenderScene:
    """The ray-cast mesh (1M-triangle displaced sphere) as a render scene:
    Lambertian, under a 2 x 2 rectangle light, dim background; `smooth`
    adds vertex normals for interpolated shading."""
    sc = RenderScene()
    v, f = displaced_sphere(levels)
    sc.meshes.append((v, f, vertex_normals(v, f)) if smooth else (v, f))
    sc.shaders = [(SHADER_LAMBERT, rec709(0.75, 0.72, 0.7), 0.0, 0.0)]
    sc.instances.append(dict(kind="mesh", index=0, shader=0, xforms=None))
    sc.background = rec709(0.03, 0.03, 0.04)
    sc.rect_lights.append((2.0, 2.0, rec709(40.0, 38.0, 33.0)))
    sc.instances.append(dict(kind="rect", index=0, shader=None,
                             xforms=np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -2.5]], np.float32)))
    sc.camera = dict(xforms=look_at((0.0, -3.2, 1.2), (0.0, 0.0, 0.0))[None], fov=float(np.float32(np.radians(45.0))),
                     aperture=0.0, focus=1.0)
    sc.settings = dict(width=width, height=height, spp=spp, max_bounces=max_bounces, seed=seed)
    return sc


# Mesh library of the large instanced scene: (kind, size, seed).  Sphere
# sizes are 4 * 4^levels triangles, heightfield patches 2 * nx * ny.
LARGE_MESHES = (
    ("patch", (72, 70), 31),     # 10,080
    ("sphere", 6, 32),           # 16,384
    ("patch", (160, 128), 33),   # 40,960
    ("sphere", 7, 34),           # 65,536
    ("patch", (256, 256), 35),   # 131,072
    ("sphere", 8, 36),           # 262,144
    ("patch", (512, 512), 37),   # 524,288
    ("sphere", 9, 38),           # 1,048,576
)


def _rock_patch(nx: int, ny: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """A heightfield of 2 * nx * ny triangles centred on the origin, about
    1.6 units across and 0.5 high."""
    v, f = heightfield(nx, ny, seed=seed, spacing=1.6 / max(nx, ny), amplitude=0.35)
    v = v.astype(np.float64)
    v[:, :2] -= (v[:, :2].max(axis=0) + v[:, :2].min(axis=0)) / 2
    v[:, 2] -= v[:, 2].min()
    return v.astype(np.float32) + np.float32(0.0), f


def large_scene(grid: int = 64, lights: int = 1024, width: int = 3840, height: int = 2160, spp: int = 128,
                max_bounces: int = 8, seed: int = 0, mesh_levels_cap: int = 9) -> RenderScene:
    """The batched large-scene workload: grid x grid instances (4096 by
    default) on a seeded jittered grid, each drawing one of eight seeded
    meshes of 10K-1M triangles (four displaced spheres, four heightfield
    patches; 2,099,040 unique triangles) with a random orientation and
    scale, and one of four material families: Lambertian, GGX metal of
    roughness 0.05-0.8, a rough dielectric of IOR 1.5 (GGX alpha 0.05; the
    reference has no transmissive closure, so this is the build's own
    PSY_SHADER_DIELECTRIC) and emissive.  Lit by `lights` blackbody sphere lights and one
    distant disk light (the sun).  `mesh_levels_cap` lowers the sphere
    meshes for small test renders."""
    sc = RenderScene()
    for kind, size, mseed in LARGE_MESHES:
        if kind == "sphere":
            v, f = displaced_sphere(min(size, mesh_levels_cap), amplitude=0.25, seed=mseed, frequency=2.0)
            v = v * np.float32(0.7)
        else:
            nx, ny = size
            if mesh_levels_cap < 9:
                nx, ny = max(8, nx >> (9 - mesh_levels_cap)), max(8, ny >> (9 - mesh_levels_cap))
            v, f = _rock_patch(nx, ny, mseed)
        sc.meshes.append((v, f))
    # Shaders: 0-5 Lambertian, 6-13 GGX metal, 14 dielectric, 15-16 emissive.
    lam = [(0.8, 0.8, 0.8), (0.7, 0.25, 0.2), (0.25, 0.6, 0.3), (0.2, 0.3, 0.7), (0.75, 0.65, 0.3),
           (0.45, 0.4, 0.35)]
    sc.shaders = [(SHADER_LAMBERT, rec709(*c), 0.0, 0.0) for c in lam]
    metals = [(1.0, 0.78, 0.34), (0.95, 0.64, 0.54), (0.91, 0.92, 0.92), (0.56, 0.57, 0.58)]
    for k in range(8):
        rough = 0.05 + 0.75 * k / 7.0
        sc.shaders.append((SHADER_GGX, rec709(*metals[k % 4]), float(np.float32(rough)), 1.0))
    sc.shaders.append((SHADER_DIELECTRIC, rec709(0.95, 0.97, 0.98), 0.05, 1.5))
    sc.shaders.append((SHADER_EMIT, (COLOR_TEMPERATURE, 3200.0, 4.0, 0.0), 0.0, 0.0))
    sc.shaders.append((SHADER_EMIT, (COLOR_TEMPERATURE, 7500.0, 4.0, 0.0), 0.0, 0.0))
    n = grid * grid
    u = uniform01(seed * 11 + 41, 10 * n).astype(np.float64).reshape(n, 10)
    spacing = 2.0
    half = grid * spacing / 2
    for i in range(n):
        gx, gy = i % grid, i // grid
        base = np.array([(gx - (grid - 1) / 2) * spacing + (u[i, 0] - 0.5) * spacing * 0.5,
                         (gy - (grid - 1) / 2) * spacing + (u[i, 1] - 0.5) * spacing * 0.5, 0.0])
        mesh = min(7, int(u[i, 2] * 8))
        rot = _rotation(_sphere_point(u[i, 3], u[i, 4]), 2.0 * np.pi * u[i, 5])
        scale = 0.6 + 0.5 * u[i, 6]
        m = u[i, 7]
        if m < 0.4:
            shader = min(5, int(u[i, 8] * 6))
        elif m < 0.7:
            shader = 6 + min(7, int(u[i, 8] * 8))
        elif m < 0.92:
            shader = 14
        else:
            shader = 15 + int(u[i, 8] < 0.5)
        if LARGE_MESHES[mesh][0] == "patch":
            rot = _rotation((0.0, 0.0, 1.0), 2.0 * np.pi * u[i, 5])
        else:
            base[2] = 0.7 * scale
        inv = rot.T / scale
        t = -inv @ base
        xf = np.concatenate([inv, t[:, None]], axis=1).astype(np.float32).ravel()
        sc.instances.append(dict(kind="mesh", index=mesh, shader=shader, xforms=xf[None]))
    lu = uniform01(seed * 11 + 42, 5 * lights).astype(np.float64).reshape(lights, 5)
    for k in range(lights):
        radius = 0.05 + 0.15 * lu[k, 0]
        temp = 2500.0 + 6500.0 * lu[k, 1]
        sc.sphere_lights.append((radius, (COLOR_TEMPERATURE, temp, 30.0, 0.0)))
        pos = np.array([(lu[k, 2] - 0.5) * 2 * half, (lu[k, 3] - 0.5) * 2 * half, 1.8 + 2.2 * lu[k, 4]])
        sc.instances.append(dict(kind="sphere", index=k, shader=None,
                                 xforms=world_to_object(np.eye(3), pos)[None]))
    sc.background = rec709(0.02, 0.025, 0.04)
    sc.distant_lights.append((0.02, (-0.3, 0.45, -0.4), (COLOR_TEMPERATURE, 5800.0, 2.0, 0.0)))
    cam_pos = np.array([-half * 0.55, -half * 0.75, half * 0.32])
    target = np.array([half * 0.05, half * 0.05, 0.0])
    sc.camera = dict(xforms=look_at(cam_pos, target)[None], fov=float(np.float32(np.radians(50.0))),
                     aperture=0.0, focus=1.0)
    sc.settings = dict(width=width, height=height, spp=spp, max_bounces=max_bounces, seed=seed)
    return sc
