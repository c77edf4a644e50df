"""GPU traversal parity: the HIP path through the C ABI against the oracle on
the same seeded rays and scenes.

Bar (from the path's specification): hit kind, primitive and instance ids
identical except for rays whose two nearest reference t-values lie within
8 ulps; t and barycentrics within 1e-6 relative.  On single-level scenes the
GPU walks every ray through the same node sequence as the reference, so the
tests there demand bit-identical records, ties included.
"""

import numpy as np
import pytest

from psychopath_amd import scenes
from psychopath_amd.tracer import Assembly, DeviceHitRecords, DeviceHits, DeviceRays, Mesh, RayBatch, Tracer

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6
TIE_ULPS = 8


def ray_batch(r):
    return RayBatch(r["ox"], r["oy"], r["oz"], r["dx"], r["dy"], r["dz"], r["time"], r["max_t"],
                    r.get("occlusion"))


def gpu_trace(tracer, r, records=False):
    h = tracer.trace(ray_batch(r), records=records)
    return dict(kind=h.kind, t=h.t, b0=h.b0, b1=h.b1, b2=h.b2, prim_id=h.prim_id, inst_id=h.inst_id)


def ulp_dist(a, b):
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def assert_exact(g, o, label=""):
    """Bit-identical hit records (kind everywhere, the rest for surface hits)."""
    bad = g["kind"] != o["kind"]
    assert not bad.any(), f"{label}: kind differs on {bad.sum()} rays, first {np.flatnonzero(bad)[:8]}"
    s = o["kind"] == 2
    for k in ("t", "b0", "b1", "b2"):
        diff = (g[k][s].view(np.uint32) != o[k][s].view(np.uint32))
        assert not diff.any(), f"{label}: {k} differs on {diff.sum()} rays"
    for k in ("prim_id", "inst_id"):
        diff = g[k][s] != o[k][s]
        assert not diff.any(), f"{label}: {k} differs on {diff.sum()} rays"


def assert_within_bar(g, o, label=""):
    """The specification's bar with the tie exemption."""
    s = (o["kind"] == 2) & (g["kind"] == 2)
    kind_bad = g["kind"] != o["kind"]
    # A kind mismatch is allowed only at a grazing/tie t (not expected).
    assert kind_bad.sum() <= max(2, int(1e-5 * len(g["kind"]))), f"{label}: kind differs on {kind_bad.sum()}"
    same = s & (g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"])
    tie = s & ~same & (ulp_dist(g["t"], o["t"]) <= TIE_ULPS)
    wrong = s & ~same & ~tie
    assert not wrong.any(), f"{label}: {wrong.sum()} rays hit different primitives outside ties"
    for k in ("t", "b0", "b1", "b2"):
        a, b = g[k][same].astype(np.float64), o[k][same].astype(np.float64)
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        ok = (err <= REL_TOL) | (np.abs(a - b) <= 1e-7)
        assert ok.all(), f"{label}: {k} beyond 1e-6 relative on {(~ok).sum()} rays"
    return int(tie.sum())


def build_pair(meshes, instances):
    from oracle.psyref import OracleScene
    a = Assembly()
    o = OracleScene()
    for v, t in meshes:
        a.add_mesh(Mesh(v, t))
        o.add_mesh(v, t)
    for m, xf in instances:
        a.add_instance(m, xf)
        o.add_instance(m, xf)
    return a.build(), o.build()


@pytest.fixture(scope="module")
def small():
    v, f = scenes.displaced_sphere(5)
    s, o = build_pair([(v, f)], [(0, None)])
    return v, f, s, o, Tracer(s, max_batch=1 << 16)


@pytest.fixture(scope="module")
def big():
    v, f = scenes.displaced_sphere(9)
    s, o = build_pair([(v, f)], [(0, None)])
    return v, f, s, o, Tracer(s, max_batch=1 << 22)


def test_small_closest_hit_exact(small):
    v, f, s, o, tr = small
    r = scenes.incoherent_rays(v.min(0), v.max(0), 50_000, seeds=(101, 102, 103))
    assert_exact(gpu_trace(tr, r), o.trace(r), "small closest")


def test_small_any_hit_exact(small):
    v, f, s, o, tr = small
    r = scenes.shadow_rays(v.min(0), v.max(0), 50_000, seeds=(104, 105))
    g, w = gpu_trace(tr, r), o.trace(r)
    assert np.array_equal(g["kind"], w["kind"])
    assert set(np.unique(g["kind"])) <= {0, 1}


def test_mixed_batch_exact(small):
    # A reference batch mixes camera/bounce (closest) and shadow (any) rays.
    v, f, s, o, tr = small
    a = scenes.incoherent_rays(v.min(0), v.max(0), 20_000, seeds=(111, 112, 113))
    b = scenes.shadow_rays(v.min(0), v.max(0), 20_000, seeds=(114, 115))
    a["occlusion"] = np.zeros(20_000, np.uint8)
    r = {k: np.concatenate([a[k], b[k]]) for k in a}
    perm = np.random.default_rng(1).permutation(40_000)
    r = {k: r[k][perm] for k in r}
    assert_exact(gpu_trace(tr, r), o.trace(r), "mixed")


def test_big_mesh_exact_on_sample(big):
    v, f, s, o, tr = big
    r = scenes.incoherent_rays(v.min(0), v.max(0), 200_000, seeds=(2, 3, 4))
    assert_exact(gpu_trace(tr, r), o.trace(r), "1M closest")


def test_big_mesh_anyhit_shadow_exact(big):
    """Any-hit shadow segments on the 1M-triangle mesh (origin and end
    uniform in the bounds, max_t 1, the benchmark's generator and seeds):
    occlusion bits identical to the oracle on a 200,000-ray sample."""
    v, f, s, o, tr = big
    r = scenes.shadow_rays(v.min(0), v.max(0), 200_000)
    g, w = gpu_trace(tr, r), o.trace(r)
    assert np.array_equal(g["kind"], w["kind"])
    assert set(np.unique(g["kind"])) <= {0, 1}
    assert 0.05 < (w["kind"] == 1).mean() < 0.95


def test_big_mesh_device_records_exact(big):
    """32-byte hit records written on the device (the benchmark's output
    form) against the oracle, and a miss record holds zeros and PSY_NONE."""
    v, f, s, o, tr = big
    r = scenes.incoherent_rays(v.min(0), v.max(0), 200_000, seeds=(12, 13, 14))
    rays = DeviceRays(ray_batch(r))
    out = DeviceHitRecords(rays.n)
    tr.trace_device(rays, out)
    rec = out.records()
    h = Hits_view(rec)
    assert_exact(h, o.trace(r), "1M closest, records")
    miss = rec["kind"] == 0
    assert miss.any()
    assert np.all(rec["t"][miss] == 0) and np.all(rec["prim_id"][miss] == 0xFFFFFFFF)
    assert np.all(rec["reserved"] == 0)


def Hits_view(rec):
    return dict(kind=rec["kind"].astype(np.uint8), t=rec["t"], b0=rec["b0"], b1=rec["b1"], b2=rec["b2"],
                prim_id=rec["prim_id"], inst_id=rec["inst_id"])


def test_hit_records_equal_arrays(small, instanced):
    """Records and arrays carry the same hits for closest, any-hit and
    instanced/motion rays, through the simple and persistent kernels and
    the host-memory path's staging."""
    v, f, s, o, tr = small
    si, oi, tri = instanced
    lo, hi = np.array([-10, -10, -2], np.float32), np.array([10, 10, 2], np.float32)
    cases = ((tr, scenes.incoherent_rays(v.min(0), v.max(0), 70_000, seeds=(171, 172, 173))),
             (tr, scenes.shadow_rays(v.min(0), v.max(0), 30_000, seeds=(174, 175))),
             (tri, scenes.incoherent_rays(lo, hi, 30_000, seeds=(176, 177, 178))))
    for tracer, r in cases:
        for shape in ((1, 1, 1, 8, 4), (0, 16, 16, 16, 4, 0), (0, 16, 32, 24, 4, 24), (2, 16, 16, 8, 4, 0),
                      (2, 12, 8, 16, 4, 16)):
            tracer.set_tuning(*shape)
            assert_exact(gpu_trace(tracer, r, records=True), gpu_trace(tracer, r), f"records {shape}")
        tracer.set_tuning(0, 16, 16, 16, 4, 0)


def _finalize_case(tracer, oracle_scene, r, label, exact=True, lights=False):
    """psy_finalize_hits on device records against the oracle's
    SurfaceIntersectionData for the same rays (light hits carry no primitive
    in the reference, so only their instance and distance are compared)."""
    rays = DeviceRays(ray_batch(r))
    recs = DeviceHitRecords(rays.n)
    tracer.trace_device(rays, recs)
    g = tracer.finalize(rays, recs)
    rec = recs.records()
    w = oracle_scene.trace(r, want_geometry=True)
    same = ((rec["kind"] == 2) & (w["kind"] == 2) & (rec["inst_id"] == w["inst_id"])
            & (rec["t"].view(np.uint32) == w["t"].view(np.uint32)))
    if not lights:
        same &= rec["prim_id"] == w["prim_id"]
        assert same.sum() > 0.2 * len(rec), label
    if exact:
        assert same.sum() == (w["kind"] == 2).sum(), label
    mine = np.concatenate([g["pos"], g["pos_err"][:, None], g["nor"], g["nor_g"], g["local_space"],
                           g["sample_pdf"][:, None], g["reserved"][:, None]], 1)
    want = w["geometry"]
    bad = (mine[same].view(np.uint32) != want[same].view(np.uint32)).any(1)
    assert not bad.any(), f"{label}: geometry differs on {bad.sum()} of {same.sum()} hits, e.g. " \
                          f"{mine[same][bad][:1]} vs {want[same][bad][:1]}"
    assert np.all(mine[rec["kind"] != 2] == 0)
    return int(same.sum())


def test_finalize_hits_match_oracle(small, instanced):
    """The hit geometry (pos, pos_err, nor, nor_g, local_space, sample_pdf)
    bit for bit against the oracle, which computes it during intersection as
    triangle_mesh.rs:264-317 does: a static mesh, instances with static and
    motion transforms and a deforming mesh, smooth-shaded meshes, and
    rectangle and sphere light instances."""
    from oracle.psyref import OracleScene
    v, f, s, o, tr = small
    _finalize_case(tr, o, scenes.incoherent_rays(v.min(0), v.max(0), 40_000, seeds=(181, 182, 183)), "static")
    si, oi, tri = instanced
    lo, hi = np.array([-10, -10, -2], np.float32), np.array([10, 10, 2], np.float32)
    _finalize_case(tri, oi, scenes.incoherent_rays(lo, hi, 40_000, seeds=(184, 185, 186)), "instanced", exact=False)
    # Smooth shading: vertex normals with two time samples under motion.
    rng = np.random.default_rng(19)
    vs, fs = scenes.displaced_sphere(4)
    vd = np.stack([vs, (vs + rng.normal(0, 0.02, size=vs.shape)).astype(np.float32)])
    nd = np.stack([scenes.vertex_normals(vd[k], fs) for k in range(2)]).astype(np.float32)
    a = Assembly()
    a.add_mesh(Mesh(vd, fs, nd))
    xfs = np.stack([rot_xf(0.1 * k, 2, (0.2 * k, 0.0, 0.0)) for k in range(3)])
    a.add_instance(0, xfs)
    os_ = OracleScene()
    os_.add_mesh(vd, fs, nd)
    os_.add_instance(0, xfs)
    os_.build()
    trs = Tracer(a.build(), max_batch=1 << 16)
    _finalize_case(trs, os_, scenes.incoherent_rays(vs.min(0) - 0.5, vs.max(0) + 0.5, 40_000, seeds=(187, 188, 189)),
                   "smooth motion", exact=False)


def test_finalize_light_hits_match_oracle():
    from types import SimpleNamespace
    from oracle import psyref
    from psychopath_amd.tracer import RectLight, SphereLight
    import psychopath_amd._lib as _lib
    rect = RectLight(1.5, 0.75, (0, 2.0, 2.0, 2.0))
    sph = SphereLight(0.6, (1, 4000.0, 1.0, 0.0))
    x_rect = np.stack([rot_xf(0.4, 0, (0.0, 0.0, -1.0))])
    x_sph = np.stack([rot_xf(0.2, 1, (1.5, 0.5, 0.0)), rot_xf(0.3, 1, (1.7, 0.4, 0.1))])
    asm = Assembly()
    asm.add_rect_light(rect)
    asm.add_sphere_light(sph)
    asm.add_instance(0, x_rect, _lib.PSY_NONE, _lib.PSY_OBJECT_RECT_LIGHT)
    asm.add_instance(0, x_sph, _lib.PSY_NONE, _lib.PSY_OBJECT_SPHERE_LIGHT)
    ns = SimpleNamespace(meshes=[], rect_lights=[(rect.width, rect.height, rect.color)],
                         sphere_lights=[(sph.radius, sph.color)],
                         instances=[dict(kind="rect", index=0, shader=None, xforms=x_rect),
                                    dict(kind="sphere", index=0, shader=None, xforms=x_sph)])
    tr = Tracer(asm.build(), max_batch=1 << 16)
    lo, hi = np.array([-3, -3, -3], np.float32), np.array([3, 3, 3], np.float32)
    n = _finalize_case(tr, psyref._assembly(ns), scenes.incoherent_rays(lo, hi, 60_000, seeds=(190, 191, 192)),
                       "lights", exact=False, lights=True)
    assert n > 1000


def test_big_mesh_coherent_exact(big):
    v, f, s, o, tr = big
    r = scenes.coherent_rays(v.min(0), v.max(0), 512, 512)
    g, w = gpu_trace(tr, r), o.trace(r)
    assert_exact(g, w, "coherent")
    assert (w["kind"] == 2).mean() > 0.2


def test_node_visit_counts_match_reference(big):
    # ACCEL_NODE_RAY_TESTS semantics (src/accel/mod.rs:22-25): the GPU's
    # per-ray internal-node visits sum to the reference's ray-node tests.
    v, f, s, o, tr = big
    r = scenes.incoherent_rays(v.min(0), v.max(0), 100_000, seeds=(7, 8, 9))
    rays = DeviceRays(ray_batch(r))
    hits = DeviceHits(rays.n)
    c = tr.trace_counted(rays, hits)
    w = o.trace(r)
    assert c["node_visits"] == w["node_tests"]
    assert c["rays"] == 100_000
    assert_exact({k: getattr(hits.to_host(), k) for k in w if k in ("kind", "t", "b0", "b1", "b2", "prim_id",
                                                                      "inst_id")}, w, "counted")


def test_full_16m_properties(big):
    # All 16,777,216 incoherent rays on the GPU: size-independent invariants,
    # determinism across runs, and a 65,536-ray sample against the oracle.
    v, f, s, o, tr = big
    n = 1 << 24
    r = scenes.incoherent_rays(v.min(0), v.max(0), n)
    rays = DeviceRays(ray_batch(r))
    h1, h2 = DeviceHits(n), DeviceHits(n)
    tr.trace_device(rays, h1)
    tr.trace_device(rays, h2)
    a, b = h1.to_host(), h2.to_host()
    assert np.array_equal(a.kind, b.kind)
    surf = a.kind == 2
    for k in ("t", "b0", "b1", "b2", "prim_id"):
        assert np.array_equal(getattr(a, k)[surf], getattr(b, k)[surf]), k
    assert set(np.unique(a.kind)) <= {0, 2}
    assert 0.3 < surf.mean() < 0.95
    assert np.all(a.t[surf] > 0)
    bsum = a.b0[surf].astype(np.float64) + a.b1[surf] + a.b2[surf]
    assert np.all(np.abs(bsum - 1.0) < 1e-5)
    assert np.all(a.prim_id[surf] < len(f))
    # The hit point lies on the hit triangle's plane (stored order v0, v2, v1).
    idx = np.flatnonzero(surf)[:: 997]
    tri = f[a.prim_id[idx]]
    p = (np.stack([r["ox"][idx], r["oy"][idx], r["oz"][idx]], 1).astype(np.float64)
         + a.t[idx, None] * np.stack([r["dx"][idx], r["dy"][idx], r["dz"][idx]], 1))
    q = (a.b0[idx, None] * v[tri[:, 0]] + a.b1[idx, None] * v[tri[:, 2]] + a.b2[idx, None] * v[tri[:, 1]])
    assert np.max(np.abs(p - q)) < 1e-4
    sample = np.random.default_rng(0).choice(n, 65_536, replace=False)
    rs = {k: r[k][sample] for k in r}
    w = o.trace(rs)
    g = dict(kind=a.kind[sample], t=a.t[sample], b0=a.b0[sample], b1=a.b1[sample], b2=a.b2[sample],
             prim_id=a.prim_id[sample], inst_id=a.inst_id[sample])
    assert_exact(g, w, "16M sample")
    assert tr.rays_traced() >= 2 * n


def test_edge_cases(small):
    v, f, s, o, tr = small
    # empty batch
    assert len(tr.trace(RayBatch([], [], [], [], [], [])).kind) == 0
    # ragged sizes around the workgroup width, axis-parallel rays with signed
    # zeros, origins on vertices, max_t = 0
    rng = np.random.default_rng(3)
    n = 1000
    o_ = v[rng.integers(0, len(v), n)]
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:200, 0] = 0.0
    d[200:400, 1] = -0.0
    d[400:500, :2] = 0.0
    d[400:500, 2] = np.where(np.arange(100) % 2, -1.0, 1.0)
    r = dict(ox=o_[:, 0], oy=o_[:, 1], oz=o_[:, 2], dx=d[:, 0], dy=d[:, 1], dz=d[:, 2],
             time=np.zeros(n, np.float32), max_t=np.where(np.arange(n) % 10 == 0, 0.0, np.inf).astype(np.float32))
    for m in (1, 63, 64, 65, 127, 129, 1000):
        sub = {k: r[k][:m] for k in r}
        assert_exact(gpu_trace(tr, sub), o.trace(sub), f"edge {m}")


def test_host_and_device_paths_agree(small):
    v, f, s, o, tr = small
    r = scenes.incoherent_rays(v.min(0), v.max(0), 70_000, seeds=(121, 122, 123))
    g = gpu_trace(tr, r)   # host path, 65,536-ray chunks
    rays = DeviceRays(ray_batch(r))
    hits = DeviceHits(rays.n)
    tr.trace_device(rays, hits)
    h = hits.to_host()
    assert np.array_equal(g["kind"], h.kind)
    s_ = g["kind"] == 2
    assert np.array_equal(g["t"][s_], h.t[s_]) and np.array_equal(g["prim_id"][s_], h.prim_id[s_])


def rot_xf(angle, axis, t):
    c, s_ = np.cos(angle), np.sin(angle)
    m = {0: [[1, 0, 0], [0, c, -s_], [0, s_, c]], 1: [[c, 0, s_], [0, 1, 0], [-s_, 0, c]],
         2: [[c, -s_, 0], [s_, c, 0], [0, 0, 1]]}[axis]
    m = np.array(m, np.float64)
    return np.array([m[0, 0], m[0, 1], m[0, 2], t[0], m[1, 0], m[1, 1], m[1, 2], t[1],
                     m[2, 0], m[2, 1], m[2, 2], t[2]], np.float32)


@pytest.fixture(scope="module")
def instanced():
    # A grid of instances: static, identity-free rotations, and 3-sample
    # motion transforms; one mesh with 2 deformation time samples.
    rng = np.random.default_rng(17)
    v, f = scenes.displaced_sphere(4)
    vd = np.stack([v, (v + rng.normal(0, 0.02, size=v.shape)).astype(np.float32)])
    inst = []
    for i in range(64):
        gx, gy = i % 8, i // 8
        base = np.array([-(gx * 2.5 - 9), -(gy * 2.5 - 9), 0.0])
        k = i % 4
        if k == 0:
            xf = None if i % 8 == 0 else np.stack([rot_xf(0.0, 0, base)])
        elif k == 1:
            xf = np.stack([rot_xf(rng.uniform(0, 0.5), int(rng.integers(3)), base)])
        else:
            xf = np.stack([rot_xf(rng.uniform(0, 0.5), int(rng.integers(3)),
                                  base + rng.uniform(-0.5, 0.5, 3)) for _ in range(3)])
        inst.append((i % 2, xf))
    s, o = build_pair([(v, f), (vd, f)], inst)
    return s, o, Tracer(s, max_batch=1 << 16)


def test_instanced_motion_within_bar(instanced):
    s, o, tr = instanced
    lo, hi = np.array([-10, -10, -2], np.float32), np.array([10, 10, 2], np.float32)
    r = scenes.incoherent_rays(lo, hi, 60_000, seeds=(131, 132, 133))
    g, w = gpu_trace(tr, r), o.trace(r)
    assert (w["kind"] == 2).mean() > 0.05
    ties = assert_within_bar(g, w, "instanced")
    assert ties <= 0.001 * len(g["kind"])
    sh = scenes.shadow_rays(lo, hi, 40_000, seeds=(134, 135))
    g, w = gpu_trace(tr, sh), o.trace(sh)
    assert (g["kind"] != w["kind"]).sum() <= 2


def test_launch_shapes_give_identical_hits(small, instanced):
    # The simple grid kernel and persistent kernels of every LDS depth and
    # refill threshold must write the same records (static scene and the
    # instanced/motion scene, which selects the other kernel specialisations).
    v, f, s, o, tr = small
    r = scenes.incoherent_rays(v.min(0), v.max(0), 30_000, seeds=(141, 142, 143))
    si, oi, tri = instanced
    lo, hi = np.array([-10, -10, -2], np.float32), np.array([10, 10, 2], np.float32)
    ri = scenes.incoherent_rays(lo, hi, 30_000, seeds=(144, 145, 146))
    # Static instances (transforms without time samples): the cooperative
    # fetch's instanced specialisation.
    vs, fs = scenes.displaced_sphere(4)
    rng = np.random.default_rng(23)
    sx, ox = build_pair([(vs, fs)], [(0, np.stack([rot_xf(rng.uniform(0, 1), k % 3, rng.uniform(-6, 6, 3))]))
                                     for k in range(24)])
    trx = Tracer(sx, max_batch=1 << 16)
    rx = scenes.incoherent_rays(np.array([-7, -7, -7], np.float32), np.array([7, 7, 7], np.float32), 30_000,
                                seeds=(147, 148, 149))
    assert_within_bar(gpu_trace(trx, rx), ox.trace(rx), "static instances")
    for tracer, rays in ((tr, r), (tri, ri), (trx, rx)):
        base = None
        for shape in ((1, 1, 1, 8, 4), (0, 1, 1, 8, 4), (0, 8, 16, 16, 8), (0, 16, 64, 24, 4), (0, 32, 32, 8, 8),
                      (0, 4, 2, 8, 8), (0, 16, 16, 16, 4, 1), (0, 24, 32, 24, 8, 8), (0, 8, 8, 8, 4, 32),
                      (2, 16, 16, 8, 4), (2, 12, 1, 16, 4), (2, 16, 32, 8, 4, 24), (2, 8, 64, 24, 4, 1)):
            tracer.set_tuning(*shape)
            g = gpu_trace(tracer, rays)
            if base is None:
                base = g
            else:
                assert_exact(g, base, f"shape {shape}")
        tracer.set_tuning(0, 16, 16, 16, 4, 0)


def test_every_ray_written_once_at_ragged_sizes(small):
    """The persistent kernel hands each ray of a launch to exactly one lane
    (waves take chunks of rays from the counter and refill lanes from them):
    with the device records pre-filled with a sentinel, every record of
    ragged batch sizes is written and equals the one-lane grid kernel's, for
    launch shapes from one wave per CU to 32, refilling at 1 to 64 idle
    lanes."""
    v, f, s, o, tr = small
    r = scenes.incoherent_rays(v.min(0), v.max(0), 40_001, seeds=(151, 152, 153))
    shapes = ((1, 1, 1, 8, 4), (0, 1, 1, 8, 4, 0), (0, 1, 64, 16, 4, 0), (0, 24, 16, 8, 4, 32),
              (0, 32, 8, 24, 8, 0), (2, 16, 16, 8, 4))
    for n in (1, 63, 65, 4097, 40_001):
        dr = DeviceRays(ray_batch({k: a[:n] for k, a in r.items()}))
        ref = None
        for shape in shapes:
            tr.set_tuning(*shape)
            hits = DeviceHitRecords(n)
            hits.buf.upload(np.full(n * 32, 0xFF, np.uint8))
            tr.trace_device(dr, hits)
            rec = hits.records()
            assert (rec["kind"] <= 2).all(), (n, shape)
            if ref is None:
                ref = rec
            else:
                assert rec.tobytes() == ref.tobytes(), (n, shape)
    tr.set_tuning(0, 16, 16, 16, 4, 0)


def test_deep_stack_spills_to_private_memory():
    # 4000 copies of one triangle build a deep tree (median splits): many
    # pending entries overflow the 8-entry LDS ring into private memory and
    # come back; hits must still match the oracle exactly.
    n = 4000
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (n, 1))
    v = (v + np.repeat(np.arange(n, dtype=np.float32)[:, None] * np.float32(1e-3), 3, axis=0)
         * np.array([0, 0, 1], np.float32)).astype(np.float32)
    f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    s, o = build_pair([(v, f)], [(0, None)])
    assert s.info()["max_depth"] >= 6
    tr = Tracer(s, max_batch=1 << 16)
    r = scenes.incoherent_rays(np.array([-0.5, -0.5, -1], np.float32), np.array([1.5, 1.5, 5], np.float32), 20_000,
                               seeds=(161, 162, 163))
    want = o.trace(r)
    for shape in ((0, 8, 8, 8, 8), (0, 8, 8, 8, 4), (1, 1, 1, 8, 4), (0, 8, 8, 8, 4, 4), (2, 8, 8, 8, 4),
                  (2, 8, 8, 8, 4, 4)):
        tr.set_tuning(*shape)
        assert_exact(gpu_trace(tr, r), want, f"deep {shape}")


def test_autotune_picks_a_valid_shape(small):
    v, f, s, o, tr = small
    r = scenes.incoherent_rays(v.min(0), v.max(0), 200_000, seeds=(151, 152, 153))
    rays = DeviceRays(ray_batch(r))
    hits = DeviceHits(rays.n)
    before = tr.rays_traced()
    best = tr.autotune(rays, hits)
    assert best == tr.tuning()
    assert best["kernel"] in (0, 1, 2) and best["stack_lds"] in (8, 16, 24) and best["simd_waves"] in (4, 8)
    assert 0 <= best["leaf_min"] <= 64
    assert tr.rays_traced() == before
    assert_exact(gpu_trace(tr, r), o.trace(r), "after autotune")
    tr.set_tuning(0, 16, 16, 16, 4, 0)


def test_time_sampled_light_sizes_match_oracle():
    """Sphere radii and rectangle dimensions as time samples: the trace
    kernels lerp them at each ray's time like the oracle, on rays grazing the
    light at every time between the samples: the same hits at bit-identical
    distances on the same instance (a light hit carries no primitive or
    barycentrics in the reference, rectangle_light.rs:300-330)."""
    from types import SimpleNamespace
    from oracle import psyref
    from psychopath_amd.tracer import RectLight, SphereLight
    import psychopath_amd._lib as _lib
    white = (0, 1.0, 1.0, 1.0)
    ident = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float32)
    n = 4096
    times = (np.arange(n, dtype=np.float32) + np.float32(0.5)) / np.float32(n)
    offs = np.linspace(0.0, 3.5, n, dtype=np.float32)[::-1].copy()
    for kind, light in (("sphere", SphereLight([1.0, 3.0, 2.0], [white, (0, 2.0, 2.0, 2.0), white])),
                        ("rect", RectLight([1.0, 3.0, 2.0], [1.0, 2.0, 4.0], [white, (0, 3.0, 1.0, 1.0)]))):
        asm = Assembly()
        if kind == "sphere":
            asm.add_sphere_light(light)
            k = _lib.PSY_OBJECT_SPHERE_LIGHT
            r = dict(ox=np.full(n, -10.0, np.float32), oy=offs, oz=np.zeros(n, np.float32),
                     dx=np.ones(n, np.float32), dy=np.zeros(n, np.float32), dz=np.zeros(n, np.float32))
            ns = SimpleNamespace(meshes=[], rect_lights=[], sphere_lights=[(light.radius, light.color)],
                                 instances=[dict(kind="sphere", index=0, shader=None, xforms=ident)])
        else:
            asm.add_rect_light(light)
            k = _lib.PSY_OBJECT_RECT_LIGHT
            r = dict(ox=offs * np.float32(0.5), oy=offs * np.float32(0.7), oz=np.full(n, 5.0, np.float32),
                     dx=np.zeros(n, np.float32), dy=np.zeros(n, np.float32), dz=np.full(n, -1.0, np.float32))
            ns = SimpleNamespace(meshes=[], rect_lights=[(light.width, light.height, light.color)], sphere_lights=[],
                                 instances=[dict(kind="rect", index=0, shader=None, xforms=ident)])
        asm.add_instance(0, ident, _lib.PSY_NONE, k)
        r["time"] = times
        r["max_t"] = np.full(n, np.inf, np.float32)
        tr = Tracer(asm.build(), max_batch=n)
        want = psyref._assembly(ns).trace(r)
        got = gpu_trace(tr, r)
        hits = int((want["kind"] != 0).sum())
        print(f"{kind} light with time samples: {hits} of {n} rays hit", flush=True)
        assert 0.1 * n < hits < 0.9 * n
        assert np.array_equal(got["kind"], want["kind"]), f"{kind}: hit kinds differ"
        s = want["kind"] != 0
        assert np.array_equal(got["t"][s].view(np.uint32), want["t"][s].view(np.uint32)), f"{kind}: t differs"
        assert np.array_equal(got["inst_id"][s], want["inst_id"][s])


def test_nested_assemblies_under_rotation_trace_exact():
    """A sub-assembly of two boxes placed by a rotating, moving instance:
    lerped boxes are not conservative under rotation, so which rays reach
    the boxes depends on walking the assembly's own BVH inside the moving
    instance's box (tracer.rs:87-150); the GPU walks the same nested BVHs as
    the oracle and returns identical records on every ray.  Also a static
    second instance of the same assembly and one without transforms."""
    import psychopath_amd._lib as _lib
    from oracle.psyref import OracleScene
    rot, w2o = scenes._rotation, scenes.world_to_object
    v1, f1 = scenes._box((0.0, 0.0, 0.0), (0.3, 0.3, 0.2), 10.0)
    v2, f2 = scenes._box((0.6, 0.2, 0.0), (0.2, 0.2, 0.4), 30.0)
    moving = np.stack([w2o(rot((0.0, 0.0, 1.0), a), np.array([-2.7, 2.4, 3.3 + 0.4 * a])) for a in (0.0, 0.6)])
    still = w2o(rot((1.0, 0.0, 0.0), 0.3), np.array([-1.5, 2.0, 3.0]))[None]
    lamp = Assembly()
    lamp.add_mesh(Mesh(v1, f1))
    lamp.add_mesh(Mesh(v2, f2))
    lamp.add_instance(0)
    lamp.add_instance(1, w2o(rot((0.0, 1.0, 0.0), 0.2), np.array([0.1, 0.0, 0.0]))[None])
    asm = Assembly()
    asm.add_mesh(Mesh(v2, f2))
    asm.add_instance(0, scenes.translate_xf((-2.0, 1.0, 2.5))[None])
    a = asm.add_assembly(lamp)
    asm.add_instance(a, moving, kind=_lib.PSY_OBJECT_ASSEMBLY)
    asm.add_instance(a, still, kind=_lib.PSY_OBJECT_ASSEMBLY)
    asm.add_instance(a, None, kind=_lib.PSY_OBJECT_ASSEMBLY)
    scene = asm.build()
    tracer = Tracer(scene, max_batch=1 << 18)

    o_lamp = OracleScene()
    o_lamp.add_mesh(v1, f1)
    o_lamp.add_mesh(v2, f2)
    o_lamp.add_instance(0)
    o_lamp.add_instance(1, w2o(rot((0.0, 1.0, 0.0), 0.2), np.array([0.1, 0.0, 0.0]))[None])
    o_lamp.build()
    o = OracleScene()
    o.add_mesh(v2, f2)
    o.add_instance(0, scenes.translate_xf((-2.0, 1.0, 2.5))[None])
    oa = o.add_assembly(o_lamp)
    o.add_assembly_instance(oa, moving)
    o.add_assembly_instance(oa, still)
    o.add_assembly_instance(oa, None)
    o.build()

    rng = np.random.default_rng(7)
    n = 200_000
    centers = np.array([[-2.7, 2.4, 3.4], [-1.5, 2.0, 3.1], [0.2, 0.1, 0.2], [-2.0, 1.0, 2.6]])
    tgt = centers[rng.integers(0, 4, n)] + rng.uniform(-0.9, 0.9, (n, 3))
    org = tgt + rng.normal(0, 1, (n, 3)) * 3
    dd = tgt - org
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    r = {k: v.astype(np.float32) for k, v in dict(ox=org[:, 0], oy=org[:, 1], oz=org[:, 2], dx=dd[:, 0],
                                                     dy=dd[:, 1], dz=dd[:, 2]).items()}
    r["time"] = rng.uniform(0, 1, n).astype(np.float32)
    r["max_t"] = np.full(n, np.inf, np.float32)
    want = o.trace(r, threads=8)
    assert (want["kind"] == 2).sum() > 10_000
    for kernel in (_lib.PSY_KERNEL_PERSISTENT, _lib.PSY_KERNEL_SIMPLE):
        tracer.set_tuning(kernel=kernel)
        got = tracer.trace(RayBatch(r["ox"], r["oy"], r["oz"], r["dx"], r["dy"], r["dz"], r["time"], r["max_t"]))
        assert np.array_equal(got.kind, want["kind"]), kernel
        s = want["kind"] == 2
        for k in ("t", "b0", "b1", "b2", "prim_id"):
            assert np.array_equal(getattr(got, k)[s], want[k][s]), (kernel, k)
        # Hit records number the flattened instances depth first (the root's
        # mesh, then each lamp's two); the oracle names an instance by its
        # index in its own assembly.
        local = np.array([0, 0, 1, 0, 1, 0, 1], np.uint32)
        assert np.array_equal(local[got.inst_id[s]], want["inst_id"][s]), kernel
