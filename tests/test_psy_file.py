""".psy reader (include/psy_file.h, csrc/psy_file.cpp) against the oracle's
restatement of the reference reader (oracle/psyfile.py), and the oracle
against the reference's own unit tests of its tokenizer and number parsers
(src/parse/data_tree.rs:565-742, src/parse/basics.rs:41-116).  Host code
only: no GPU needed."""

import math
import random

import numpy as np
import pytest

from oracle import psyfile as opf
from psychopath_amd import psyfile, scenes

GOLDEN_EXAMPLE = "tests/golden/example_cornell_box.psy"


# ---------------------------------------------------------------- oracle pins

def tok(text, pos=0):
    kind, val, end = opf.next_token(text.encode(), pos)
    return kind, val, end


def test_oracle_tokenizer_matches_reference_unit_tests():
    # data_tree.rs tokenize_1 .. tokenize_7
    assert tok("Thing") == ("TypeName", "Thing", 5)
    assert tok("  \n# gdfgdf gfdg dggdf\\sg dfgsd \n   Thing") == ("TypeName", "Thing", 41)
    t = " Thing { }"
    k1 = tok(t)
    k2 = tok(t, k1[2])
    k3 = tok(t, k2[2])
    assert (k1[0], t[k1[2]:]) == ("TypeName", " { }")
    assert (k2[0], t[k2[2]:]) == ("OpenInner", " }")
    assert (k3[0], t[k3[2]:]) == ("CloseInner", "")
    assert tok(" $hi_there ") == ("Ident", "$hi_there", 10)
    assert tok(" $hi\\ t\\#he\\[re ") == ("Ident", "$hi\\ t\\#he\\[re", 15)
    t = " $hi the[re"
    seq, p = [], 0
    for _ in range(5):
        k, v, p = tok(t, p)
        seq.append((k, v, p))
    assert seq == [("Ident", "$hi", 4), ("TypeName", "the", 8), ("OpenLeaf", None, 9),
                   ("TypeName", "re", 11), ("End", None, 11)]
    t = "Thing $yar { # A comment\n\tThing2 []\n}"
    seq, p = [], 0
    for _ in range(8):
        k, v, p = tok(t, p)
        seq.append((k, p))
    assert seq == [("TypeName", 5), ("Ident", 10), ("OpenInner", 12), ("TypeName", 32), ("OpenLeaf", 34),
                   ("CloseLeaf", 35), ("CloseInner", 37), ("End", 37)]


def test_oracle_tree_and_iterators_match_reference_unit_tests():
    # parse_1
    root = opf.parse_data_tree("\n            Thing {}\n        ")
    assert root["type"] == "ROOT" and root["ident"] is None and len(root["children"]) == 1
    c = root["children"][0]
    assert c["internal"] and c["type"] == "Thing" and c["ident"] is None and c["children"] == []
    # iter_1 .. iter_3
    root = opf.parse_data_tree("A {}\nB {}\nA []\nA {}\nB {}")
    assert sum(1 for c in root["children"] if c["type"] == "A") == 3
    assert sum(1 for c in root["children"] if c["type"] == "A" and c["internal"]) == 2
    root = opf.parse_data_tree("A []\nB {}\nA {}\nA []\nB {}")
    assert sum(1 for c in root["children"] if c["type"] == "A" and not c["internal"]) == 2


def test_oracle_number_parsers_match_reference_unit_tests():
    for s in ("42", "     42", "42   "):
        assert opf.ws_uint(s) == (42, len(s))
    assert opf.ws_uint("     42   53") == (42, 10)
    for s, v in (("42", 42.0), ("     42", 42.0), ("42   ", 42.0), ("42.5", 42.5), ("     42.5", 42.5),
                 ("42.5   ", 42.5), ("-42.5", -42.5), ("     -42.5", -42.5), ("-42.5   ", -42.5)):
        assert opf.ws_f32(s) == (v, len(s))
    assert opf.ws_f32("     42   53") == (42.0, 10)
    assert opf.ws_f32("     -42.5   53") == (-42.5, 13)
    # ws_f32_4: errors
    assert opf.ws_f32("a1.0") is None
    assert opf._all_f32("0abc", 1) is None
    assert opf._tuple_f32("0.abc 1.2", 2) is None
    # ws_i32 style negatives are not u32
    assert opf.ws_uint("-42") is None


def test_oracle_float_rounding_closed_forms():
    """Decimal -> f32 rounding: exact values, halfway cases (round half to
    even), overflow and the smallest subnormal."""
    assert opf._round_f32("0.5") == 0.5
    assert opf._round_f32("16777217") == 16777216.0          # halfway, ties to even
    assert opf._round_f32("16777219") == 16777220.0          # halfway, ties to even (up)
    assert opf._round_f32("1e39") == math.inf
    assert opf._round_f32("1.4e-45") == 2.0 ** -149
    rng = random.Random(5)
    for _ in range(2000):
        x = rng.uniform(-1e6, 1e6) * 10 ** rng.randint(-30, 30)
        s = "%.12g" % x
        assert opf._round_f32(s) == float(np.float32(np.float64(s))) or abs(
            opf._round_f32(s) - float(np.float32(np.float64(s)))) <= abs(float(np.float32(x))) * 2 ** -23


def test_rec709_matrix_matches_build_script_closed_form():
    """The equal-energy Rec.709 matrix maps RGB (1,1,1) to XYZ (1,1,1) and
    its rows are the primaries' chromaticities scaled (build.rs:178-225)."""
    m = opf.rec709_e_matrix().astype(np.float64)
    assert np.allclose(m.sum(axis=1), 1.0, atol=1e-7)
    for col, (x, y) in enumerate(((0.64, 0.33), (0.30, 0.60), (0.15, 0.06))):
        s = m[:, col].sum()
        assert abs(m[0, col] / s - x) < 1e-6 and abs(m[1, col] / s - y) < 1e-6


# ------------------------------------------------------------- C++ vs oracle

def oracle_tree(text):
    return opf.dump_tree(opf.parse_data_tree(text))


REFERENCE_TEXTS = [
    "Thing", "  \n# gdfgdf gfdg dggdf\\sg dfgsd \n   Thing {}", " Thing { }", "A { $hi_there [x] }",
    "Thing $yar { # A comment\n\tThing2 []\n}", "\n            Thing {}\n        ",
    "A {}\nB {}\nA []\nA {}\nB {}", "A []\nB {}\nA {}\nA []\nB {}",
    "X $a\\ b\\#c\\[d { Y [1 2 \\] 3] }", "Unicodé $naïve { Lëaf [ünï] }", "# only a comment", "",
    "A { B { C { D [deep] } } }\r\nE [x]\r\n",
]


def outcome_cpp(text):
    try:
        return ("ok", psyfile.tree(text))
    except psyfile.PsyFileError as e:
        return ("err", e.kind, e.offset)


def outcome_oracle(text):
    try:
        return ("ok", oracle_tree(text))
    except opf.DataTreeError as e:
        return ("err", e.kind, e.offset)


@pytest.mark.parametrize("text", REFERENCE_TEXTS)
def test_tree_matches_oracle(text):
    assert outcome_cpp(text) == outcome_oracle(text)


BAD_TEXTS = [
    ("Thing $x [", "MissingOpenInternal"), ("Thing {", "MissingCloseInternal"), ("Thing [abc", "MissingCloseLeaf"),
    ("Thing ]", "MissingOpener"), ("Thing $a $b", "MissingOpenInternal"), ("}", "Other"), ("A {} \\", "Other"),
    ("A { [x] }", "MissingCloseInternal"), ("A [x\\]", "MissingCloseLeaf"),
]


@pytest.mark.parametrize("text,kind", BAD_TEXTS)
def test_tree_errors_match_oracle(text, kind):
    with pytest.raises(opf.DataTreeError) as oe:
        opf.parse_data_tree(text)
    with pytest.raises(psyfile.PsyFileError) as ce:
        psyfile.tree(text)
    assert oe.value.kind == kind
    assert (ce.value.kind, ce.value.offset) == (oe.value.kind, oe.value.offset)


def test_tree_fuzz_matches_oracle():
    """Random edits of a real scene text: both readers accept or reject
    alike, with the same tree or the same error kind and offset."""
    base = psyfile.write(scenes.cornell_box(8, 8, 1, 1, 0))
    rng = random.Random(11)
    alphabet = "{}[]$#\\ \n\tab0.-e"
    for _ in range(400):
        t = list(base)
        for _ in range(rng.randint(1, 4)):
            i = rng.randrange(len(t))
            op = rng.random()
            if op < 0.4:
                del t[i]
            elif op < 0.8:
                t.insert(i, rng.choice(alphabet))
            else:
                t[i] = rng.choice(alphabet)
        text = "".join(t)
        assert outcome_cpp(text) == outcome_oracle(text), text


def oracle_scene(parsed):
    """The oracle's parse in the descriptors' terms."""
    asm = parsed["assembly"]
    kinds = {"Lambert": scenes.SHADER_LAMBERT, "GGX": scenes.SHADER_GGX, "Emit": scenes.SHADER_EMIT}
    shaders = [(kinds[k], opf.color_desc(c), r, f) for k, c, r, f in asm["shaders"]]
    meshes, rects, spheres, where = [], [], [], []
    for kind, o in asm["objects"]:
        if kind == "mesh":
            where.append(("mesh", len(meshes)))
            meshes.append(o)
        elif kind == "rect":
            where.append(("rect", len(rects)))
            rects.append(([d[0] for d in o["dims"]], [d[1] for d in o["dims"]],
                          [opf.color_desc(c) for c in o["color"]]))
        else:
            where.append(("sphere", len(spheres)))
            spheres.append((o["radius"], [opf.color_desc(c) for c in o["color"]]))
    insts = []
    for i in asm["instances"]:
        kind, idx = where[i["object"]]
        xf = i["xforms"]
        if not xf and kind != "mesh":
            xf = [[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]]
        insts.append((kind, idx, i["shader"], np.asarray(xf, np.float32).reshape(-1, 12)))
    cam = parsed["camera"]
    world = parsed["world"]
    distant = [(d["radius"], [tuple(v) for v in d["direction"]], [opf.color_desc(c) for c in d["color"]])
               for d in world["distant_lights"]]
    return dict(shaders=shaders, meshes=meshes, rects=rects, spheres=spheres, instances=insts,
                camera=(np.asarray(cam["xforms"], np.float32), cam["fovs"], cam["apertures"], cam["focus"]),
                background=opf.color_desc(world["background"]), distant=distant,
                settings=(parsed["width"], parsed["height"], parsed["spp"], parsed["seed"]))


def f32eq(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def assert_same_scene(sc, o):
    assert (sc.settings["width"], sc.settings["height"], sc.settings["spp"], sc.settings["seed"]) == o["settings"]
    assert sc.settings["max_bounces"] == 2
    assert f32eq(sc.camera["xforms"], o["camera"][0])
    for k, want in zip(("fov", "aperture", "focus"), o["camera"][1:]):
        assert f32eq(np.atleast_1d(sc.camera[k]), want)
    assert sc.background[0] == o["background"][0] and f32eq(sc.background[1:], o["background"][1:])
    assert len(sc.shaders) == len(o["shaders"])
    for a, b in zip(sc.shaders, o["shaders"]):
        assert a[0] == b[0] and a[1][0] == b[1][0] and f32eq(a[1][1:], b[1][1:]) and f32eq(a[2:], b[2:])
    assert len(sc.meshes) == len(o["meshes"])
    for m, om in zip(sc.meshes, o["meshes"]):
        assert f32eq(m[0], np.asarray(om["verts"], np.float32))
        assert np.array_equal(m[1], np.asarray(om["tris"], np.uint32).reshape(-1, 3))
        if om["normals"] is None:
            assert len(m) == 2
        else:
            assert f32eq(m[2], np.asarray(om["normals"], np.float32))
    def colours(c):  # a colour or a list of time samples, as kinds and values
        cs = c if isinstance(c, list) else [c]
        return [x[0] for x in cs], [x[1:] for x in cs]

    def same_colours(a, b):
        (ka, va), (kb, vb) = colours(a), colours(b)
        return ka == kb and f32eq(va, vb)

    assert len(sc.rect_lights) == len(o["rects"]) and len(sc.sphere_lights) == len(o["spheres"])
    for a, b in zip(sc.rect_lights, o["rects"]):
        assert f32eq(np.atleast_1d(a[0]), b[0]) and f32eq(np.atleast_1d(a[1]), b[1]) and same_colours(a[2], b[2])
    for a, b in zip(sc.sphere_lights, o["spheres"]):
        assert f32eq(np.atleast_1d(a[0]), b[0]) and same_colours(a[1], b[1])
    assert len(sc.instances) == len(o["instances"])
    for a, (kind, idx, shader, xf) in zip(sc.instances, o["instances"]):
        assert (a["kind"], a["index"], a["shader"]) == (kind, idx, shader)
        if a["xforms"] is None:
            assert xf.size == 0
        else:
            assert f32eq(a["xforms"], xf)
    assert len(sc.distant_lights) == len(o["distant"])
    for a, b in zip(sc.distant_lights, o["distant"]):
        dirs = a[1] if isinstance(a[1], list) else [a[1]]
        assert f32eq(np.atleast_1d(a[0]), b[0]) and f32eq(dirs, b[1]) and same_colours(a[2], b[2])


def psy_expressible(sc):
    """The scene with its dielectric shaders (which the .psy format cannot
    hold) replaced by GGX ones."""
    sc.shaders = [(scenes.SHADER_GGX, c, r, 0.04) if k == scenes.SHADER_DIELECTRIC else (k, c, r, f)
                  for k, c, r, f in sc.shaders]
    return sc


def generated_scenes():
    return [
        ("cornell", scenes.cornell_box()),
        ("motion", scenes.motion_instances(grid=2, levels=2, lights=3, width=64, height=36, spp=2)),
        ("terrain", scenes.terrain_scene(nx=16, ny=12, width=32, height=32, spp=2)),
        ("large", psy_expressible(scenes.large_scene(grid=3, lights=5, width=32, height=18, spp=1,
                                                     mesh_levels_cap=2))),
    ]


def test_dielectric_shaders_are_refused_by_the_writer():
    with pytest.raises(ValueError, match="only Lambert, GGX and Emit"):
        psyfile.write(scenes.large_scene(grid=3, lights=5, width=32, height=18, spp=1, mesh_levels_cap=2))


@pytest.mark.parametrize("label,sc", generated_scenes())
def test_reader_matches_oracle_on_generated_scenes(label, sc):
    text = psyfile.write(sc, label, f"{label}.png")
    got = psyfile.read(text)
    assert len(got) == 1 and got[0].name == label and got[0].output_path == f"{label}.png"
    assert_same_scene(got[0], oracle_scene(opf.parse_file(text)))


@pytest.mark.parametrize("label,sc", generated_scenes())
def test_written_scene_reads_back(label, sc):
    """Transforms, vertices, indices, light sizes, settings and the camera
    read back bit for bit; colours within the f32 rounding of the Rec.709
    conversion."""
    got = psyfile.read(psyfile.write(sc))[0]
    assert len(got.meshes) == len(sc.meshes)
    for a, b in zip(got.meshes, sc.meshes):
        v = np.asarray(b[0], np.float32)
        assert f32eq(a[0], v if v.ndim == 3 else v[None])
        assert np.array_equal(a[1], np.asarray(b[1], np.uint32).reshape(-1, 3))
    for a, b in zip(got.instances, sc.instances):
        assert (a["kind"], a["index"], a["shader"]) == (b["kind"], b["index"], b.get("shader"))
        if b.get("xforms") is not None:
            assert f32eq(a["xforms"], np.asarray(b["xforms"], np.float32).reshape(-1, 12))
    assert f32eq(got.camera["xforms"], np.asarray(sc.camera["xforms"], np.float32).reshape(-1, 12))
    assert abs(got.camera["fov"] - sc.camera["fov"]) <= 2e-7 * abs(sc.camera["fov"])
    for a, b in zip(got.shaders, sc.shaders):
        assert a[0] == b[0] and a[1][0] == b[1][0]
        assert np.allclose(a[1][1:], b[1][1:], rtol=3e-6, atol=1e-7)


def test_example_scene_is_rejected_like_the_reference():
    """The reference's example Cornell box keeps shaders and objects outside
    the root Assembly, so its first Instance names missing data: parse_scene
    fails with InstancedMissingData at that Data leaf (psy_assembly.rs:77-87)."""
    text = open(GOLDEN_EXAMPLE, encoding="utf-8").read()
    assert psyfile.tree(text) == oracle_tree(text)
    with pytest.raises(opf.PsyParseError) as oe:
        opf.parse_file(text)
    with pytest.raises(psyfile.PsyFileError) as ce:
        psyfile.read(text)
    assert oe.value.kind == ce.value.kind == "InstancedMissingData"
    assert ce.value.offset == oe.value.offset == text.encode().index(b"Data [$__Area]") + len("Data")


SCENE_ERRORS = [
    # (edit of a valid file, expected kind)
    (lambda t: t.replace("    Output {\n        Path [\"render.png\"]\n    }\n", ""), "WrongNodeCount"),
    (lambda t: t.replace('Path ["render.png"]', "Path [render.png]"), "IncorrectLeafData"),
    (lambda t: t.replace("SamplesPerPixel [1]", "SamplesPerPixel [x]"), "IncorrectLeafData"),
    (lambda t: t.replace("SamplesPerPixel [1]\n", ""), "MissingNode"),
    (lambda t: t.replace("Resolution [8 8]", "Resolution [8]"), "IncorrectLeafData"),
    (lambda t: t.replace("Type [Color]", "Type [Sky]"), "UnknownVariant"),
    (lambda t: t.replace("SurfaceShaderBind [$shader0]", "SurfaceShaderBind [$nope]", 1), "UnknownVariant"),
    (lambda t: t.replace("Data [$mesh0]", "Data [$nope]"), "InstancedMissingData"),
    (lambda t: t.replace("Type [Lambert]", "Type [Glass]", 1), "UnknownVariant"),
    (lambda t: t.replace("FaceVertCounts [3 3]", "FaceVertCounts [2 4]", 1), "IncorrectLeafData"),
    (lambda t: t.replace("FaceVertCounts [3 3]", "FaceVertCounts [3 3 3]", 1), "IncorrectLeafData"),
    (lambda t: t.replace("Assembly {", "Assembly {\n        Assembly {\n        }", 1), "UnknownError"),
    (lambda t: t.replace("Assembly {", "Assembly {\n        Instance {\n            Data [$later]\n        }\n"
                         "        Assembly $later {\n        }", 1), "InstancedMissingData"),
]


@pytest.mark.parametrize("edit,kind", SCENE_ERRORS)
def test_scene_errors_match_oracle(edit, kind):
    text = edit(psyfile.write(scenes.cornell_box(8, 8, 1, 1, 0)))
    with pytest.raises(opf.PsyParseError) as oe:
        opf.parse_file(text)
    with pytest.raises(psyfile.PsyFileError) as ce:
        psyfile.read(text)
    assert oe.value.kind == kind
    assert ce.value.kind == kind


def test_polygons_fan_and_numbers_in_leaves():
    """Fan triangulation of a pentagon (psy_mesh_surface.rs:96-115), numbers
    with exponents, signs and surrounding whitespace, a comment between
    nodes, two scenes in one file, and shader-less instances."""
    text = """# two scenes
Scene $a {
    Output { Path [ "a.exr" ] }
    RenderSettings { Resolution [ 4 2 ] SamplesPerPixel [3] }
    Camera { Fov [90] Transform [1 0 0 0 0 1 0 0 0 0 1 0 0 0 -5 1] }
    World { BackgroundShader { Type [ Color ] Color [blackbody, 5000 2.5e-1] } }
    Assembly {
        MeshSurface $p {
            Vertices [0 0 0  1 0 0  1.5 1 0  .5 2 0  -5E-1 1 0 ]
            FaceVertCounts [5]
            FaceVertIndices [0 1 2 3 4]
        }
        Instance { Data [$p] }
    }
}
Scene $b {
    Output { Path ["b.png"] }
    RenderSettings { Resolution [1 1] SamplesPerPixel [1] Seed [7] }
    Camera { Fov [45] FocalDistance [0] ApertureRadius [1] Transform [1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1] }
    World { BackgroundShader { Type [Color] Color [color_temperature, 6500 1] } }
    Assembly { }
}
"""
    got = psyfile.read(text)
    parsed = [opf.parse_scene(c) for c in opf.parse_data_tree(text)["children"] if c["type"] == "Scene"]
    assert len(got) == 2
    for g, p in zip(got, parsed):
        assert_same_scene(g, oracle_scene(p))
    a = got[0]
    assert np.array_equal(a.meshes[0][1], [[0, 1, 2], [0, 2, 3], [0, 3, 4]])
    assert f32eq(a.meshes[0][0][0, 3:], [[0.5, 2, 0], [-0.5, 1, 0]])
    assert a.instances[0]["shader"] is None and a.instances[0]["xforms"] is None
    assert a.settings["seed"] == 0 and a.background == (scenes.COLOR_BLACKBODY, 5000.0, 0.25, 0.0)
    b = got[1]
    # focus distance 0 disables the lens (camera.rs:52-58)
    assert (b.camera["aperture"], b.camera["focus"]) == (0.0, 1.0) and b.settings["seed"] == 7


def test_light_time_samples_are_read():
    """Repeated Dimensions, Radius, Direction and Color leaves are time
    samples (psy_light.rs:16-184), kept in file order; colour samples must
    share one representation, as Lerp for Color requires (color.rs:331)."""
    sc = scenes.motion_instances(grid=2, levels=2, lights=2, width=16, height=16, spp=1)
    sc.rect_lights.append(([1.0, 2.0, 0.5], [1.5, 1.0, 2.5], [(0, 1.0, 2.0, 3.0), (0, 3.0, 2.0, 1.0)]))
    sc.sphere_lights[0] = ([0.25, 0.5], sc.sphere_lights[0][1])
    sc.sphere_lights[1] = (0.3, [(2, 4000.0, 30.0, 0.0), (2, 6500.0, 60.0, 0.0), (2, 5000.0, 45.0, 0.0)])
    sc.distant_lights.append(([0.02, 0.04], [(-0.4, 0.3, -0.6), (0.3, -0.4, -0.7)],
                              [(2, 5800.0, 100.0, 0.0), (2, 5900.0, 120.0, 0.0)]))
    sc.instances.append(dict(kind="rect", index=len(sc.rect_lights) - 1, shader=None,
                             xforms=np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -3]], np.float32)))
    text = psyfile.write(sc)
    got = psyfile.read(text)[0]
    assert_same_scene(got, oracle_scene(opf.parse_file(text)))
    assert len(got.rect_lights[-1][0]) == 3 and len(got.rect_lights[-1][2]) == 2
    assert len(got.sphere_lights[0][0]) == 2 and len(got.sphere_lights[1][1]) == 3
    assert len(got.distant_lights[0][1]) == 2
    mixed = text.replace("Color [color_temperature, 6500", "Color [blackbody, 6500", 1)
    assert mixed != text
    with pytest.raises(psyfile.PsyFileError) as e:
        psyfile.read(mixed)
    assert e.value.kind == "UnknownVariant" and "different representations" in str(e.value)


def test_camera_lens_time_samples_are_read():
    """Several Fov, FocalDistance and ApertureRadius leaves are time
    samples (parse_camera, psy.rs:355-420), kept in file order; a zero focus
    distance among them, or a missing list, disables the lens (camera.rs:
    32-57)."""
    base = psyfile.write(scenes.cornell_box(8, 8, 1, 1, 0))
    head, tail = base.split("        Fov [", 1)
    tail = tail.split("\n", 1)[1]
    cam = ("        Fov [30]\n        Fov [45]\n        Fov [60]\n"
           "        FocalDistance [2]\n        FocalDistance [3]\n"
           "        ApertureRadius [0.5]\n")
    lines = [ln for ln in tail.split("\n", 2)[:2]]
    assert lines[0].lstrip().startswith("FocalDistance") and lines[1].lstrip().startswith("ApertureRadius")
    text = head + cam + tail.split("\n", 2)[2]
    sc = psyfile.read(text)[0]
    k = np.float32(np.float32(np.pi) / np.float32(180.0))
    assert f32eq(sc.camera["fov"], [np.float32(30) * k, np.float32(45) * k, np.float32(60) * k])
    assert f32eq(sc.camera["focus"], [2.0, 3.0]) and sc.camera["aperture"] == 0.5
    assert_same_scene(sc, oracle_scene(opf.parse_file(text)))
    # A zero focus sample disables the lens; so does a missing aperture.
    for t2 in (text.replace("FocalDistance [3]", "FocalDistance [0]"),
               text.replace("        ApertureRadius [0.5]\n", "")):
        sc2 = psyfile.read(t2)[0]
        assert (sc2.camera["aperture"], sc2.camera["focus"]) == (0.0, 1.0)
        assert_same_scene(sc2, oracle_scene(opf.parse_file(t2)))
    # The writer keeps every sample.
    again = psyfile.read(psyfile.write(sc))[0]
    assert len(again.camera["fov"]) == 3 and f32eq(again.camera["focus"], [2.0, 3.0])


def test_deep_nesting_is_refused_not_crashed():
    text = "A { " * 5000 + "} " * 5000
    with pytest.raises(psyfile.PsyFileError) as e:
        psyfile.tree(text)
    assert e.value.kind == "NestingTooDeep"
    assert psyfile.tree("A { " * 100 + "} " * 100).count("\n") == 101


# ------------------------------------------------------- nested assemblies

def _xf_text(xf):
    from psychopath_amd.psyfile import _matrix
    return _matrix(np.asarray(xf, np.float32))


def nested_cornell_text(w=32, h=24, spp=2, seed=0):
    """The Cornell box with its two boxes in a sub-assembly (own shader), the
    sub-assembly instanced twice (once plain, once moved with two time
    samples) and a second level holding one more box under a rotation."""
    from psychopath_amd.psyfile import write
    sc = scenes.cornell_box(w, h, spp, 2, seed)
    boxes = sc.meshes[5:7]
    sc.meshes = sc.meshes[:5]
    sc.instances = [i for i in sc.instances if not (i["kind"] == "mesh" and i["index"] >= 5)]
    text = write(sc, "nested", "nested.png")
    from psychopath_amd.psyfile import _g
    box_txt = []
    for k, (v, f) in enumerate(boxes):
        box_txt.append(f"            MeshSurface $box{k} {{\n                Vertices [" +
                       " ".join(_g(x) for x in np.asarray(v, np.float32).ravel()) + "]\n"
                       "                FaceVertCounts [" + " ".join("3" for _ in range(len(f))) + "]\n"
                       "                FaceVertIndices [" + " ".join(str(int(x)) for x in np.asarray(f).ravel()) +
                       "]\n            }\n")
    ident = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    move0 = [1, 0, 0, 0.4, 0, 1, 0, -0.2, 0, 0, 1, 0]
    move1 = [1, 0, 0, 0.7, 0, 1, 0, -0.5, 0, 0, 1, 0.1]
    c, s_ = np.cos(0.3), np.sin(0.3)
    rot = [c, -s_, 0, 2.8, s_, c, 0, -1.0, 0, 0, 1, 0]
    sub = ("        Assembly $boxes {\n"
           "            SurfaceShader $white2 {\n                Type [Lambert]\n"
           "                Color [rec709, 0.6 0.65 0.7]\n            }\n"
           "            SurfaceShader $gloss {\n                Type [GGX]\n"
           "                Color [rec709, 0.9 0.8 0.5]\n                Roughness [0.3]\n"
           "                Fresnel [0.8]\n            }\n"
           + "".join(box_txt) +
           "            Assembly $inner {\n"
           "                SurfaceShader $red2 {\n                    Type [Lambert]\n"
           "                    Color [rec709, 0.7 0.2 0.1]\n                }\n"
           + box_txt[1].replace("$box1", "$tiny").replace("            ", "                ", 1) +
           "                Instance {\n                    Data [$tiny]\n"
           "                    SurfaceShaderBind [$red2]\n"
           f"                    Transform [{_xf_text(rot)}]\n                }}\n"
           "            }\n"
           "            Instance {\n                Data [$box0]\n                SurfaceShaderBind [$white2]\n"
           "            }\n"
           "            Instance {\n                Data [$box1]\n                SurfaceShaderBind [$gloss]\n"
           f"                Transform [{_xf_text(ident)}]\n            }}\n"
           "            Instance {\n                Data [$inner]\n"
           f"                Transform [{_xf_text(move0)}]\n            }}\n"
           "        }\n"
           "        Instance {\n            Data [$boxes]\n        }\n"
           "        Instance {\n            Data [$boxes]\n"
           f"            Transform [{_xf_text(move0)}]\n            Transform [{_xf_text(move1)}]\n        }}\n")
    at = text.index("        Instance {")
    return text[:at] + sub + text[at:]


def same_assembly(a, b):
    """The reader's assembly `a` (a RenderScene or SubAssembly) equals the
    oracle reader's `b`: meshes, lights, instances with their own transforms
    bit for bit, and the sub-assemblies, nested as in the file."""
    assert len(a.meshes) == len(b.meshes)
    for x, y in zip(a.meshes, b.meshes):
        assert f32eq(x[0], np.asarray(y[0], np.float32).reshape(x[0].shape)) and np.array_equal(x[1], y[1])
    assert len(a.rect_lights) == len(b.rect_lights) and len(a.sphere_lights) == len(b.sphere_lights)
    for x, y in zip(a.rect_lights, b.rect_lights):
        assert f32eq(np.atleast_1d(x[0]), np.atleast_1d(y[0])) and f32eq(np.atleast_1d(x[1]), np.atleast_1d(y[1]))
    for x, y in zip(a.sphere_lights, b.sphere_lights):
        assert f32eq(np.atleast_1d(x[0]), np.atleast_1d(y[0]))
    assert len(a.instances) == len(b.instances)
    for x, y in zip(a.instances, b.instances):
        assert (x["kind"], x["index"], x["shader"]) == (y["kind"], y["index"], y["shader"])
        if y["xforms"] is None:
            assert x["xforms"] is None
        else:
            assert f32eq(x["xforms"], y["xforms"]), (x["xforms"], y["xforms"])
    assert len(a.sub_assemblies) == len(b.sub_assemblies)
    for x, y in zip(a.sub_assemblies, b.sub_assemblies):
        same_assembly(x, y)


def test_nested_assemblies_read_like_the_oracle():
    """Sub-assemblies stay nested (AssemblyBuilder::add_assembly,
    assembly.rs:197-282): the traversal walks them nested, as
    trace_assembly does (tracer.rs:87-150), with the transforms the
    transform stack composes."""
    from psychopath_amd import render
    text = nested_cornell_text()
    got = psyfile.read(text)[0]
    sc = opf.to_scene(opf.parse_file(text))
    same_assembly(got, sc)
    assert len(got.shaders) == len(sc.shaders) == 3 + 2 + 1
    assert len(got.sub_assemblies) == 1 and len(got.sub_assemblies[0].sub_assemblies) == 1
    # the moved copy of the sub-assembly has two transform samples
    assert max(0 if i["xforms"] is None else len(i["xforms"]) for i in got.instances) == 2
    render.assembly_of(got).build()  # the packer accepts it


def test_lights_inside_nested_assemblies_are_read_like_the_oracle():
    from psychopath_amd import render
    text = nested_cornell_text().replace(
        "        Assembly $boxes {\n",
        "        Assembly $boxes {\n            SphereLight $bulb {\n                Radius [0.1]\n"
        "                Color [rec709, 1 1 1]\n            }\n"
        "            Instance {\n                Data [$bulb]\n"
        f"                Transform [{_xf_text([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -1])}]\n            }}\n", 1)
    got = psyfile.read(text)[0]
    sc = opf.to_scene(opf.parse_file(text))
    same_assembly(got, sc)
    assert len(got.sub_assemblies[0].sphere_lights) == 1
    render.assembly_of(got).build()
