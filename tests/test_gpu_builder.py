"""The device BVH builder (psy_scene_create_on_device) against the host
builder, whose trees the oracle pins (tests/test_builder_parity.py): the
same BVH4s bit for bit (node structure, traversal codes, time-sampled
bounds, leaf object order), on spheres up to the 1M-triangle mesh, random
soups, forced median splits, deformation motion, the 5.2M-triangle terrain,
and the fallback to the host builder for boxes with negative zeros."""

import time

import numpy as np
import pytest

from psychopath_amd import scenes
from psychopath_amd.tracer import Assembly, Mesh

pytestmark = pytest.mark.gpu


def build_both(verts, tris):
    a = Assembly()
    a.add_mesh(Mesh(verts, tris))
    a.add_instance(0)
    t0 = time.perf_counter()
    host = a.build()
    t1 = time.perf_counter()
    dev = a.build(device=True)
    t2 = time.perf_counter()
    return host, dev, t1 - t0, t2 - t1


def assert_same_trees(host, dev):
    for m in (0, -1):
        d1, o1 = host.dump_bvh(m)
        d2, o2 = dev.dump_bvh(m)
        assert d1[:3].tolist() == d2[:3].tolist(), (m, d1[:3], d2[:3])
        np.testing.assert_array_equal(d1, d2, err_msg=f"tree {m}")
        np.testing.assert_array_equal(o1, o2, err_msg=f"order {m}")
    hi, di = host.info(), dev.info()
    for k in ("node_count", "leaf_count", "motion_bounds", "max_depth", "device_bytes"):
        assert hi[k] == di[k], k


@pytest.mark.parametrize("levels", [0, 1, 2, 3, 5, 7, 9])
def test_device_build_spheres(levels):
    v, f = scenes.displaced_sphere(levels)
    host, dev, th, td = build_both(v, f)
    assert dev.info()["device_built_meshes"] == 1
    assert_same_trees(host, dev)
    print(f"\n{len(f)} triangles: host {th * 1e3:.1f} ms, device {td * 1e3:.1f} ms "
          f"(device builder {dev.info()['device_build_ms']:.1f} ms)")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 17, 64, 65, 1000, 100_000])
def test_device_build_random_soups(n):
    rng = np.random.default_rng(n)
    v = rng.uniform(0.5, 3, size=(3 * n, 3)).astype(np.float32)
    v[: 3 * n // 2] *= -1.0
    f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    host, dev, _, _ = build_both(v, f)
    assert dev.info()["device_built_meshes"] == 1
    assert_same_trees(host, dev)


def test_device_build_forced_median_splits():
    # 2000 copies of one triangle: every SAH cost is NaN, splits peel one
    # triangle at a time and the depth guard switches to median splits.
    n = 2000
    v = np.tile(np.array([[1, 1, 1], [2, 1, 1], [1, 2, 1]], np.float32), (n, 1))
    f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    host, dev, _, _ = build_both(v, f)
    assert dev.info()["device_built_meshes"] == 1
    assert_same_trees(host, dev)
    assert host.info()["max_depth"] >= 15


@pytest.mark.parametrize("samples", [2, 3])
def test_device_build_deformation_time_samples(samples):
    rng = np.random.default_rng(samples)
    v0, f = scenes.displaced_sphere(6)
    vs = [v0]
    for k in range(1, samples):
        vs.append((v0 + rng.normal(0, 0.02 * k, size=v0.shape)).astype(np.float32) + np.float32(0.0))
    v = np.stack(vs)
    v[-1, : len(v0) // 3] += np.float32(0.5)
    host, dev, _, _ = build_both(v, f)
    assert dev.info()["device_built_meshes"] == 1
    assert_same_trees(host, dev)
    assert host.info()["motion_bounds"] > 0


def test_negative_zeros_fall_back_to_the_host_builder():
    # Triangles lying in the plane x = -0: their boxes and centroids keep the
    # negative zero, whose unions depend on order.
    v = np.array([[-0.0, 0, 0], [-0.0, 1, 0], [-0.0, 0, 1], [2, 2, 2], [3, 2, 2], [2, 3, 2]] * 4, np.float32)
    f = np.arange(24, dtype=np.uint32).reshape(8, 3)
    host, dev, _, _ = build_both(v, f)
    assert dev.info()["device_built_meshes"] == 0
    assert_same_trees(host, dev)


def test_device_build_terrain_5m():
    v, f = scenes.heightfield(2048, 1280)
    host, dev, th, td = build_both(v, f)
    assert dev.info()["device_built_meshes"] == 1
    assert_same_trees(host, dev)
    print(f"\nterrain {len(f)} triangles: host scene {th:.2f} s, device scene {td:.2f} s "
          f"(device builder {dev.info()['device_build_ms']:.1f} ms)")


def test_device_light_tree_matches_host():
    """The light trees of the large scene (1,024 sphere lights, SAH splits
    only, light_tree.rs:211-290) partitioned on the device: the renderer's
    light selection then gives the same image as with the host-built
    trees; the scene's BVHs too are the host builder's."""
    from psychopath_amd import render
    sc = scenes.large_scene(width=96, height=54, spp=2, grid=6, lights=1024, mesh_levels_cap=3)
    asm = render.assembly_of(sc)
    host = asm.build()
    dev = asm.build(device=True, device_min_tris=2)
    assert dev.info()["device_light_trees"] >= 1
    for m in range(len(sc.meshes)):
        d1, o1 = host.dump_bvh(m)
        d2, o2 = dev.dump_bvh(m)
        assert np.array_equal(d1, d2) and np.array_equal(o1, o2), m
    r1 = render.Renderer(host, sc.shaders, sc.background, sc.distant_lights)
    r2 = render.Renderer(dev, sc.shaders, sc.background, sc.distant_lights)
    a, sa = r1.render(sc.camera, sc.settings)
    b, _ = r2.render(sc.camera, sc.settings)
    # A lit image with bounces, so that equality says something.
    assert np.all(np.isfinite(a)) and a.mean() > 0
    assert sa["rays"] > sc.settings["width"] * sc.settings["height"] * sc.settings["spp"]
    assert np.array_equal(a, b)
