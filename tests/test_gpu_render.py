"""GPU wavefront renderer against the oracle's restatement of the reference
renderer (src/renderer.rs, src/scene, src/light, src/shading): same scenes,
same sample sequence, XYZ images compared per pixel.

Tolerance: a pixel matches when every channel is within 1e-3 relative of
the oracle's (plus 1e-9 absolute for black pixels); at least 99.9% of the
pixels must match.  Paths can leave the oracle's only where a transcendental
function of the device library rounds differently from the host's and that
flips a discrete choice (a triangle edge, a light-tree branch), which the
0.1% allowance covers.
"""

import numpy as np
import pytest

from oracle import psyref
from psychopath_amd import render, scenes

pytestmark = pytest.mark.gpu

REL = 1e-3
FRACTION = 0.999


def matching_fraction(img, ref):
    ok = np.abs(img.astype(np.float64) - ref) <= REL * np.abs(ref.astype(np.float64)) + 1e-9
    return ok.all(axis=-1).mean()


def gpu_render(sc, **kw):
    r = render.build(sc, **kw)
    img, stats = r.render(sc.camera, sc.settings)
    return img, stats, r


def check_parity(sc, label):
    ref, info = psyref.render(sc)
    img, stats, _ = gpu_render(sc)
    # Rays traced (RenderStats::ray_count): camera, bounce and light rays.
    assert stats["rays"] == info["rays"], (stats["rays"], info["rays"])
    frac = matching_fraction(img, ref)
    print(f"{label}: {frac * 100:.3f}% pixels within {REL} relative; mean xyz gpu {img.mean((0, 1))} "
          f"oracle {ref.mean((0, 1))}; {stats}", flush=True)
    assert np.all(np.isfinite(img))
    assert frac >= FRACTION, f"{label}: only {frac * 100:.3f}% pixels match"
    return img, ref


def test_cornell_box_matches_oracle():
    check_parity(scenes.cornell_box(96, 96, 8, 4, 0), "cornell 96x96x8")


def test_dielectric_boxes_match_oracle():
    """The rough dielectric (PSY_SHADER_DIELECTRIC, no reference counterpart;
    the oracle's restatement is pinned in tests/test_oracle_dielectric.py)
    on the Cornell box's two boxes, one clear and nearly smooth, one tinted,
    rough and denser: the GPU render matches the oracle's, and light now
    passes through the boxes.  Bad parameters are refused."""
    sc = scenes.cornell_box(96, 96, 8, 6, 0)
    sc.shaders.append((scenes.SHADER_DIELECTRIC, scenes.rec709(0.95, 0.95, 0.95), 0.02, 1.5))
    sc.shaders.append((scenes.SHADER_DIELECTRIC, scenes.rec709(0.9, 0.6, 0.4), 0.3, 1.7))
    assert sc.instances[5]["kind"] == "mesh" and sc.instances[6]["kind"] == "mesh"
    sc.instances[5]["shader"] = 3
    sc.instances[6]["shader"] = 4
    img, _ = check_parity(sc, "cornell, dielectric boxes, 96x96x8 depth 6")
    lam, _, _ = gpu_render(scenes.cornell_box(96, 96, 8, 6, 0))
    assert not np.allclose(img, lam)
    for bad in ((scenes.SHADER_DIELECTRIC, scenes.rec709(1, 1, 1), 0.0, 1.5),
                (scenes.SHADER_DIELECTRIC, scenes.rec709(1, 1, 1), 0.1, 1.0)):
        sc.shaders[3] = bad
        with pytest.raises(Exception):
            render.build(sc)


def test_cornell_box_seed_and_depth():
    check_parity(scenes.cornell_box(64, 48, 4, 2, 7), "cornell 64x48x4 depth 2 seed 7")


def test_execution_options_do_not_change_results():
    """Graph replay vs direct launches, material sort vs queue order, one
    wave vs many small waves, and the one-launch plain renderer give
    bit-identical images."""
    sc = scenes.cornell_box(64, 64, 4, 4, 0)
    base, _, r = gpu_render(sc)
    r.set_options(graphs=False)
    a, _ = r.render(sc.camera, sc.settings)
    r.set_options(graphs=True, sort=False)
    b, _ = r.render(sc.camera, sc.settings)
    r.set_options(graphs=True, sort=True)
    again, _ = r.render(sc.camera, sc.settings)
    r.set_options(plain=True)
    plain, _ = r.render(sc.camera, sc.settings)
    budgets = []
    for w in (1, 2, 3, 4):
        r.set_options(shade_waves=w)
        budgets.append(r.render(sc.camera, sc.settings)[0])
        r.set_options(plain=True, shade_waves=w)
        budgets.append(r.render(sc.camera, sc.settings)[0])
    from psychopath_amd._lib import PsyError
    with pytest.raises(PsyError):
        r.set_options(shade_waves=5)
    small, stats, _ = gpu_render(sc, max_paths=3000)
    assert stats["waves"] > 1
    for other in (a, b, again, plain, small, *budgets):
        assert np.array_equal(base, other)


def test_plain_single_launch_matches_wavefront():
    """The one-launch renderer gives the wavefront renderer's image on a
    scene with motion, instancing, sphere lights, GGX and depth of field."""
    sc = scenes.motion_instances(grid=3, levels=3, lights=3, width=48, height=27, spp=4, max_bounces=5)
    img, _, r = gpu_render(sc)
    r.set_options(plain=True)
    plain, pstats = r.render(sc.camera, sc.settings)
    assert np.array_equal(img, plain)
    _, wstats = r.render(sc.camera, sc.settings)
    r.set_options(plain=False)
    _, wstats = r.render(sc.camera, sc.settings)
    assert pstats["rays"] == wstats["rays"] > sc.settings["width"] * sc.settings["height"] * sc.settings["spp"]


def test_motion_blur_instancing_sphere_lights_dof():
    sc = scenes.motion_instances(grid=4, levels=4, lights=4, width=80, height=45, spp=8, max_bounces=6)
    check_parity(sc, "motion/instances 80x45x8")


def test_distant_light_and_emissive_surface():
    sc = scenes.cornell_box(48, 48, 4, 3, 1)
    sc.distant_lights.append((0.05, (0.3, 0.2, -1.0), (scenes.COLOR_BLACKBODY, 5800.0, 1e-6, 0.0)))
    sc.background = scenes.rec709(0.02, 0.03, 0.05)
    sc.shaders.append((scenes.SHADER_EMIT, scenes.rec709(2.0, 2.0, 2.0), 0.0, 0.0))
    sc.shaders.append((scenes.SHADER_GGX, scenes.rec709(0.9, 0.9, 0.9), 0.3, 0.5))
    sc.instances[5]["shader"] = 3   # tall box emits
    sc.instances[6]["shader"] = 4   # short box is glossy
    check_parity(sc, "cornell + distant light + emit + GGX")


def test_rect_light_furnace():
    """The oracle's closed-form scene (grey plane under a large two-sided
    rectangle light) on the GPU."""
    sc = scenes.RenderScene()
    sc.shaders = [(scenes.SHADER_LAMBERT, scenes.rec709(0.5, 0.5, 0.5), 0.0, 0.0)]
    h = 1000.0
    v = np.array([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], np.float32)
    sc.add_mesh_instance(v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), 0)
    sc.rect_lights.append((4.0, 4.0, scenes.rec709(16.0, 16.0, 16.0)))
    sc.instances.append(dict(kind="rect", index=0, shader=None,
                             xforms=np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -1.0]], np.float32)))
    sc.camera = dict(xforms=np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0.5]], np.float32), fov=1.2,
                     aperture=0.0, focus=1.0)
    sc.settings = dict(width=32, height=32, spp=16, max_bounces=3, seed=0)
    check_parity(sc, "rect furnace")


def test_full_resolution_cornell_properties():
    """At the workload's full size (512x512, 16 spp, depth 4): finite,
    non-negative Y, deterministic across calls, and the same image through
    the graph and direct launch paths."""
    sc = scenes.cornell_box()
    img, stats, r = gpu_render(sc)
    assert stats["paths"] == 512 * 512 * 16
    assert np.all(np.isfinite(img))
    assert (img[..., 1] >= 0).all()
    img2, _ = r.render(sc.camera, sc.settings)
    assert np.array_equal(img, img2)
    r.set_options(graphs=False)
    img3, _ = r.render(sc.camera, sc.settings)
    assert np.array_equal(img, img3)


def test_served_requests_match_renders_alone():
    """Requests of different sizes, spp, bounce limits and scenes interleaved
    in shared waves give the images their plain renders give."""
    cornell = scenes.cornell_box(64, 64, 4, 4, 0)
    motion = scenes.motion_instances(grid=3, levels=3, lights=3, width=48, height=32, spp=4, max_bounces=3)
    ra = render.build(cornell, max_paths=5000)
    rb = render.build(motion, max_paths=5000)
    reqs = [
        (0, cornell.camera, dict(width=64, height=64, spp=4, max_bounces=4, seed=0)),
        (1, motion.camera, dict(width=48, height=32, spp=4, max_bounces=3, seed=0)),
        (0, cornell.camera, dict(width=40, height=30, spp=9, max_bounces=2, seed=3)),
        (0, cornell.camera, dict(width=17, height=13, spp=1, max_bounces=5, seed=1)),
        (1, motion.camera, dict(width=33, height=21, spp=7, max_bounces=6, seed=2)),
    ]
    alone = [(ra if k == 0 else rb).render(cam, st)[0] for k, cam, st in reqs]
    server = render.Server([ra, rb])
    ids = [server.submit(k, cam, st) for k, cam, st in reqs]
    waves = []
    for rid, ref in zip(ids, alone):
        img, stats = server.wait(rid)
        waves.append(stats["waves"])
        assert np.array_equal(img, ref)
        assert stats["latency_ms"] >= stats["queued_ms"] >= 0
    st = server.stats()
    server.close()
    # the small path pools force the larger requests over several waves
    assert max(waves) > 1 and st["paths"] == sum(s["width"] * s["height"] * s["spp"] for _, _, s in reqs)


def test_server_budget_admission():
    """A budget that fits one film at a time still completes every request,
    in order, with the same images; a request that can never fit fails."""
    cornell = scenes.cornell_box(32, 32, 2, 2, 0)
    r = render.build(cornell, max_paths=4096)
    settings = [dict(width=32, height=32, spp=2, max_bounces=2, seed=s) for s in range(4)]
    refs = [r.render(cornell.camera, st)[0] for st in settings]
    probe = render.Server([r])
    fixed = probe.stats()["fixed_bytes"]
    probe.close()
    film = 32 * 32 * 3 * 4
    server = render.Server([r], hbm_budget=fixed + film + film // 2)
    ids = [server.submit(0, cornell.camera, st) for st in settings]
    for rid, ref in zip(ids, refs):
        img, _ = server.wait(rid)
        assert np.array_equal(img, ref)
    from psychopath_amd._lib import PsyError
    with pytest.raises(PsyError):
        server.submit(0, cornell.camera, dict(width=256, height=256, spp=1, max_bounces=1, seed=0))
    server.close()


def test_server_film_allocation_failure_fails_the_request():
    """A film allocation that fails while nothing else runs ends the request
    with PSY_E_NOMEM instead of waiting for memory that cannot free up; one
    that fails while another request renders is retried and succeeds; the
    server still serves afterwards and closes with requests queued."""
    from psychopath_amd._lib import PSY_E_NOMEM, PsyError
    sc = scenes.cornell_box(32, 32, 2, 2, 0)
    r = render.build(sc, max_paths=1024)
    ref, _ = r.render(sc.camera, sc.settings)
    server = render.Server([r])
    server.fail_film_allocs(1)
    rid = server.submit(0, sc.camera, sc.settings)
    with pytest.raises(PsyError) as e:
        server.wait(rid)
    assert e.value.code == PSY_E_NOMEM
    big = dict(sc.settings, width=96, height=96, spp=16)
    first = server.submit(0, sc.camera, big)
    while server.progress(first) == 0:
        pass
    server.fail_film_allocs(1)
    second = server.submit(0, sc.camera, sc.settings)
    img, _ = server.wait(second)
    assert np.array_equal(img, ref)
    server.wait(first, keep_image=False)
    for s in range(3):
        server.submit(0, sc.camera, dict(sc.settings, seed=s))
    server.close()


def test_autotune_keeps_images():
    sc = scenes.cornell_box(48, 48, 4, 3, 0)
    r = render.build(sc)
    before, _ = r.render(sc.camera, sc.settings)
    choice = r.autotune(sc.camera, dict(width=32, height=32, spp=2, max_bounces=3, seed=0))
    assert choice["waves_per_cu"] in (16, 24, 32) and choice["stack_lds"] in (8, 16, 24)
    assert choice["shade_waves"] in (1, 2, 3, 4) and choice["plain"] in (True, False)
    after, _ = r.render(sc.camera, sc.settings)
    assert np.array_equal(before, after)


def test_large_instanced_scene_reduced():
    """The large-scene workload's structure (eight meshes, Lambertian, GGX
    metal, rough dielectric and emissive instances, sphere lights and a sun)
    at reduced size: 256 instances of meshes capped at 65,536 triangles, 64
    lights, depth 8."""
    sc = scenes.large_scene(grid=16, lights=64, width=96, height=54, spp=8, mesh_levels_cap=7)
    check_parity(sc, "large scene reduced 96x54x8")


def test_large_instanced_scene_full_geometry():
    """The large-scene workload's full geometry and lighting (4096 instances
    over 2,099,040 unique triangles, 1024 sphere lights and a sun, depth 8)
    at a small image; the plain one-launch renderer gives the same image."""
    sc = scenes.large_scene(width=64, height=36, spp=2)
    img, _ = check_parity(sc, "large scene full 64x36x2")
    r = render.build(sc)
    r.set_options(plain=True)
    plain, _ = r.render(sc.camera, sc.settings)
    assert np.array_equal(img, plain)


def test_scene_file_renders_like_the_oracle():
    """A Cornell box written as a .psy file (parser layout), read by the
    library's reader and rendered on the GPU, against the oracle's reader
    and renderer on the same text; the reader's scene gives the same image
    as the scene it was written from when the colours read back exactly."""
    from oracle import psyfile as opf
    from psychopath_amd import psyfile
    text = psyfile.write(scenes.cornell_box(64, 48, 8, 2, 3), "cornell")
    sc = psyfile.read(text)[0]
    assert sc.settings == dict(width=64, height=48, spp=8, max_bounces=2, seed=3)
    ref, _ = psyref.render(opf.to_scene(opf.parse_file(text)))
    img, stats, _ = gpu_render(sc)
    frac = matching_fraction(img, ref)
    print(f"cornell .psy 64x48x8: {frac * 100:.3f}% pixels within {REL}; {stats}", flush=True)
    assert frac >= FRACTION


def test_smooth_shading_normals():
    """Interpolated vertex normals (triangle_mesh.rs:272-296): a static
    smooth sphere and motion-blurred instances whose normals have two time
    samples."""
    check_parity(scenes.sphere_scene(levels=4, width=48, height=48, spp=8, max_bounces=3, smooth=True),
                 "smooth sphere 48x48x8")
    check_parity(scenes.motion_instances(grid=3, levels=3, lights=3, width=48, height=27, spp=4, max_bounces=4,
                                         smooth=True), "smooth motion instances 48x27x4")


def test_mesh_without_shader_emits_magenta():
    """A mesh instance without a shader renders as the reference's magenta
    emitter (tracer.rs:161-164)."""
    sc = scenes.cornell_box(48, 48, 4, 3, 2)
    sc.instances[5]["shader"] = None   # the tall box
    img, _ = check_parity(sc, "cornell, tall box without a shader")
    shaded, _, _ = gpu_render(scenes.cornell_box(48, 48, 4, 3, 2))
    # the emitting box adds light, more in X and Z (magenta) than in Y
    gain = img.mean((0, 1)) - shaded.mean((0, 1))
    assert gain[0] > gain[1] > 0 and gain[2] > gain[1]


def test_served_rows_stream_as_they_finish():
    """A request rendered over many waves delivers its finished rows while
    later rows still render; the rows put together equal the image the
    request gives when collected whole, and equal a plain render."""
    sc = scenes.cornell_box(64, 48, 8, 3, 5)
    r = render.build(sc, max_paths=2048)
    alone, _ = r.render(sc.camera, sc.settings)
    server = render.Server([r])
    rid = server.submit(0, sc.camera, sc.settings)
    parts, seen = [], []
    for row, rows in server.stream(rid):
        seen.append(row)
        parts.append(rows)
    server.close()
    img = np.concatenate(parts, axis=0)
    assert img.shape == alone.shape and np.array_equal(img, alone)
    assert len(parts) > 1 and seen == sorted(seen)
    from psychopath_amd._lib import PsyError
    server = render.Server([r])
    rid = server.submit(0, sc.camera, sc.settings)
    with pytest.raises(PsyError):
        server.read(rid, 0, 64 * 48 + 1)
    server.wait(rid)
    server.close()


def test_command_line_renders_a_scene_file(tmp_path, capsys):
    """`python -m psychopath_amd -i scene.psy` writes the scene's Output
    path as Image::write_png does; --serialized_output streams the same
    pixels in the bucket protocol; --crop keeps the rectangle."""
    from psychopath_amd import __main__ as cli
    from psychopath_amd import film, psyfile
    out = tmp_path / "cornell.png"
    text = psyfile.write(scenes.cornell_box(40, 30, 2, 2, 0), "cornell", str(out))
    src = tmp_path / "cornell.psy"
    src.write_text(text)
    assert cli.main(["-i", str(src), "-s", "4"]) == 0
    sc = psyfile.read(text)[0]
    sc.settings["spp"] = 4
    img, _, _ = gpu_render(sc)
    assert out.read_bytes() == film.encode(img, film.FORMATS[".png"])
    capsys.readouterr()
    assert cli.main(["-i", str(src), "-s", "4", "--serialized_output"]) == 0
    lines = capsys.readouterr().out.splitlines()
    assert lines[0] == "DIV" and lines[-2:] == ["BUCKET_END", "DIV"]
    assert lines[3] == film.bucket_base64(img, 0, 0, 32, 30)
    assert cli.main(["-i", str(src), "-s", "4", "--crop", "5", "6", "20", "10"]) == 0
    cropped = np.zeros_like(img)
    cropped[6:11, 5:21] = img[6:11, 5:21]
    assert out.read_bytes() == film.encode(cropped, film.FORMATS[".png"])


def test_crop_region_renders_the_same_pixels():
    """A crop region renders only its pixels, leaving the rest black, with
    the values of the full render (both forms, and served with rows
    streamed from the region)."""
    sc = scenes.cornell_box(48, 40, 4, 3, 1)
    full, stats, r = gpu_render(sc)
    crop = dict(sc.settings, crop=(7, 5, 20, 13))
    want = np.zeros_like(full)
    want[5:18, 7:27] = full[5:18, 7:27]
    img, cstats = r.render(sc.camera, crop)
    assert np.array_equal(img, want) and cstats["paths"] == 20 * 13 * 4
    r.set_options(plain=True)
    img, _ = r.render(sc.camera, crop)
    assert np.array_equal(img, want)
    r.set_options(plain=False)
    small = render.build(sc, max_paths=512)
    server = render.Server([small])
    rid = server.submit(0, sc.camera, crop)
    rows = np.concatenate([part for _, part in server.stream(rid)], axis=0)
    server.close()
    assert np.array_equal(rows, full[5:18, 7:27])
    from psychopath_amd._lib import PsyError
    with pytest.raises(PsyError):
        r.render(sc.camera, dict(sc.settings, crop=(40, 0, 10, 5)))


def test_nested_assemblies_render_like_the_oracle():
    """A scene file with two levels of sub-assemblies (one instanced twice,
    once under motion): the reader flattens it with the transform stack's
    arithmetic; the oracle renders it nested, walking the sub-assemblies as
    tracer.rs does."""
    import importlib
    from oracle import psyfile as opf
    from psychopath_amd import psyfile
    tp = importlib.import_module("tests.test_psy_file")
    text = tp.nested_cornell_text(48, 36, 8, 1)
    sc = psyfile.read(text)[0]
    ref, info = psyref.render(opf.to_scene(opf.parse_file(text)))
    img, stats, _ = gpu_render(sc)
    frac = matching_fraction(img, ref)
    print(f"nested cornell 48x36x8: {frac * 100:.3f}% pixels within {REL}; rays {stats['rays']} vs {info['rays']}",
          flush=True)
    assert frac >= FRACTION


def test_nested_assembly_lights_match_oracle():
    """Lights inside nested assemblies (lamps, each with a bulb sub-assembly,
    one lamp under motion): the packer flattens the geometry and keeps one
    light tree per assembly, and light selection recurses through them as
    Assembly::sample_lights does (scene/assembly.rs:46-141), against the
    oracle's restatement of that recursion; the one-launch form and a served
    request give the same image."""
    sc = scenes.nested_lights_scene(80, 80, 8, 3, 0)
    img, ref = check_parity(sc, "nested lights 80x80x8")
    r = render.build(sc)
    r.set_options(plain=True)
    plain, _ = r.render(sc.camera, sc.settings)
    assert np.array_equal(plain, img)
    sv = render.Server([r])
    served, _ = sv.wait(sv.submit(0, sc.camera, sc.settings))
    sv.close()
    assert np.array_equal(served, img)
    # The lamps light the room: brighter than the box alone.
    base, _, _ = gpu_render(scenes.cornell_box(80, 80, 8, 3, 0))
    assert img.mean() > 1.5 * base.mean()


def test_camera_lens_time_samples_match_oracle():
    """Field of view, aperture and focus distance as time samples of
    different lengths (Camera::generate_ray lerps each list on its own,
    camera.rs:74-79): the wavefront, the one-launch form and a served
    request agree with the oracle."""
    sc = scenes.cornell_box(64, 48, 8, 3, 5)
    f = sc.camera["fov"]
    sc.camera = dict(sc.camera, fov=[f * 0.8, f, f * 1.3], aperture=[0.0, 0.35], focus=[7.0, 9.0, 8.0, 10.5])
    img, ref = check_parity(sc, "cornell 64x48x8 lens time samples")
    r = render.build(sc)
    r.set_options(plain=True)
    plain, _ = r.render(sc.camera, sc.settings)
    assert np.array_equal(plain, img)
    sv = render.Server([r])
    served, _ = sv.wait(sv.submit(0, sc.camera, sc.settings))
    sv.close()
    assert np.array_equal(served, img)
    # One sample in a list is the scalar.
    one = dict(sc.camera, fov=[f], aperture=[0.0], focus=[1.0])
    a, _ = r.render(one, sc.settings)
    b, _ = r.render(dict(sc.camera, fov=f, aperture=0.0, focus=1.0), sc.settings)
    assert np.array_equal(a, b)


def test_time_sampled_lights_match_oracle():
    """Light sizes, colours and directions as time samples (PsychoBlend
    writes one per motion-blur step): a Cornell box whose light grows and
    changes colour over the shutter, with a distant light that turns, and
    the motion workload with pulsing blackbody spheres, against the oracle."""
    sc = scenes.cornell_box(64, 48, 8, 3, 2)
    w, h, c = sc.rect_lights[0]
    sc.rect_lights[0] = ([w, 2 * w, 1.5 * w], [h, 0.5 * h, h], [c, (c[0], 0.5 * c[1], 1.5 * c[2], c[3])])
    sc.distant_lights.append(([0.05, 0.1], [(-0.4, 0.3, -0.6), (0.4, 0.3, -0.6)],
                              [(2, 5800.0, 3.0, 0.0), (2, 7000.0, 5.0, 0.0)]))
    check_parity(sc, "cornell 64x48x8 time-sampled lights")
    mo = scenes.motion_instances(grid=3, levels=3, lights=3, width=64, height=36, spp=8, max_bounces=4)
    r0, c0 = mo.sphere_lights[0]
    mo.sphere_lights[0] = ([r0, 2.5 * r0], [c0, (c0[0], c0[1] * 1.3, c0[2] * 0.4, c0[3])])
    r1, c1 = mo.sphere_lights[1]
    mo.sphere_lights[1] = (r1, [c1, (c1[0], c1[1] * 0.7, c1[2] * 2.0, c1[3]), c1])
    check_parity(mo, "motion 64x36x8 time-sampled sphere lights")


def test_motion_blurred_export_renders_like_the_oracle(tmp_path):
    """A scene file shaped like a motion-blurred PsychoBlend export (one
    Transform, Fov, FocalDistance and ApertureRadius per blur step, lights
    with per-step sizes, colours and directions) read by the product's and
    the oracle's readers and rendered by both, then written by the command
    line exactly as the rendered image encodes."""
    from oracle import psyfile as opf
    from psychopath_amd import __main__ as cli
    from psychopath_amd import film, psyfile
    sc = scenes.cornell_box(48, 36, 8, 3, 4)
    xf = np.asarray(sc.camera["xforms"], np.float32).reshape(-1, 12)
    steps = np.stack([xf[0] + np.float32(k) * np.array([0, 0, 0, 0.05, 0, 0, 0, 0.0, 0, 0, 0, 0.02], np.float32)
                      for k in range(3)])
    f = sc.camera["fov"]
    sc.camera = dict(xforms=steps, fov=[f, f * 1.05, f * 1.1], aperture=[0.02, 0.03], focus=[8.0, 8.5])
    w, h, c = sc.rect_lights[0]
    sc.rect_lights[0] = ([w, 1.2 * w, 1.4 * w], [h, h, 1.3 * h], [c, (c[0], 1.1 * c[1], c[2], 0.9 * c[3])])
    sc.distant_lights.append(([0.03, 0.04, 0.05], [(-0.4, 0.3, -0.6), (-0.3, 0.35, -0.6), (-0.2, 0.4, -0.6)],
                              (2, 6000.0, 2.0, 0.0)))
    out = tmp_path / "blur.png"
    text = psyfile.write(sc, "blur", str(out))
    got = psyfile.read(text)[0]
    assert len(got.camera["fov"]) == 3 and len(got.distant_lights[0][1]) == 3
    img, stats, _ = gpu_render(got)
    ref, info = psyref.render(opf.to_scene(opf.parse_file(text)))
    frac = matching_fraction(img, ref)
    print(f"motion-blurred export 48x36x8: {frac * 100:.3f}% pixels within {REL}; rays {stats['rays']} vs "
          f"{info['rays']}", flush=True)
    assert frac >= FRACTION and stats["rays"] == info["rays"]
    src = tmp_path / "blur.psy"
    src.write_text(text)
    assert cli.main(["-i", str(src)]) == 0
    assert out.read_bytes() == film.encode(img, film.FORMATS[".png"])
