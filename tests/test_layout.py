"""Scene layout (anchors, placement methods), imports and the Room object.

The layout cases restate the reference's own tests with their inputs and expected values
(crates/forgedthoughts/src/lib.rs:1561-1771); the import cases restate lib.rs:2078-2300 on
files written to a temporary directory, with a small library of our own standing in for
the reference's built-in one."""
import os

import numpy as np
import pytest

from forgedthoughts_amd import abi
from forgedthoughts_amd.forge.evaluator import ForgeEvalError, load_scene_source
from forgedthoughts_amd.forge.imports import ImportError_, load_program_with_imports, load_scene_file
from forgedthoughts_amd.scene import compile_scene

f32 = np.float32


def _vec(st, name, fld="pos"):
    o = st.bindings[name].value.fields[fld]
    return [o.fields[k] for k in "xyz"]


def test_attach_align_and_offset():  # lib.rs:1561-1582
    st = load_scene_source("let floor = Box { size: vec3(10.0, 0.5, 10.0) };\n"
                           "var placed = Sphere { radius: 1.0 }.attach(floor, Top).align_x(floor, Center)"
                           ".align_z(floor, Center).offset_x(0.75);")
    assert _vec(st, "placed") == [f32(0.75), f32(1.25), f32(0.0)]


def test_relative_placement_and_rotation():  # lib.rs:1585-1606
    st = load_scene_source("let anchor = Sphere { radius: 0.5 };\n"
                           "var cube = Box { size: vec3(1.0) }.right_of(anchor, 0.25).rotate_z(15.0);")
    assert _vec(st, "cube")[0] == f32(1.25)
    assert _vec(st, "cube", "rot")[2] == f32(15.0)


def test_face_to():  # lib.rs:1609-1630
    st = load_scene_source("var target = Sphere { radius: 0.5 }; target.pos.x = 1.0; target.pos.y = 1.0;\n"
                           "var box = Box { size: vec3(1.0) }.face_to(target);")
    rot = _vec(st, "box", "rot")
    assert rot[0] == f32(-45.0) and rot[1] == f32(90.0)


def test_anchor_offsets():  # lib.rs:1661-1679
    st = load_scene_source("let floor = Box { size: vec3(10.0, 0.5, 10.0) };\n"
                           "var placed = Sphere { radius: 1.0 }.attach(floor, Top + 0.1).align_z(floor, Center - 0.4);")
    p = _vec(st, "placed")
    assert p[1] == f32(1.35) and p[2] == f32(-0.4)


def test_corner_anchors():  # lib.rs:1682-1704
    st = load_scene_source("let room = Box { size: vec3(10.0, 4.0, 8.0) };\n"
                           "var cupboard = Box { size: vec3(2.0, 3.0, 1.5) }.attach(room, BackRightCorner);")
    assert _vec(st, "cupboard") == [f32(4.0), f32(-0.5), f32(-3.25)]


def test_custom_anchor_names():  # lib.rs:1707-1735
    st = load_scene_source(
        "let character = Box { size: vec3(2.0, 4.0, 1.0), anchors: { FootLeft: vec3(-0.4, -2.0, 0.0) } };\n"
        "var shoe = Box { size: vec3(0.8, 0.4, 1.0), anchors: { Mount: vec3(0.0, -0.2, 0.0) } }\n"
        "  .attach(character, \"FootLeft\", \"Mount\");")
    assert _vec(st, "shoe") == [f32(-0.4), f32(-1.8), f32(0.0)]


def test_asset_defined_anchors():  # lib.rs:1738-1771
    st = load_scene_source(
        "sdf Cabinet { let width = 2.0; let height = 3.0; let depth = 1.0;\n"
        "  let anchors = { TopSurface: vec3(0.0, height * 0.5, 0.0) };\n"
        "  fn bounds() { return vec3(width * 0.5, height * 0.5, depth * 0.5); }\n"
        "  fn distance(p) { return length(p) - 1.0; } };\n"
        "var cabinet = Cabinet {};\n"
        "var vase = Sphere { radius: 0.2 }.attach(cabinet, \"TopSurface\", Bottom);")
    assert _vec(st, "vase")[1] == f32(1.7)


def test_layout_leaves_the_receiver_unchanged_and_needs_bounds():
    st = load_scene_source("let floor = Box { size: vec3(10.0, 0.5, 10.0) };\n"
                           "let s = Sphere { radius: 1.0 };\n"
                           "let t = s.attach(floor, Top);")
    assert "pos" not in st.bindings["s"].value.fields
    assert _vec(st, "t")[1] == f32(1.25)
    with pytest.raises(ForgeEvalError):
        load_scene_source("let a = Foo { };\nlet b = Sphere { radius: 1.0 }.attach(a, Top);")


def test_layout_bounds_of_composites_and_modifiers():
    """object_bounds (eval.rs:4784-4916): unions, subtractions, mirror, repeat and slice."""
    st = load_scene_source(
        "var a = Sphere { radius: 1.0 }; a.pos.x = 2.0;\n"
        "let u = a + Sphere { radius: 0.5 };\n"
        "let m = a.mirror_x();\n"
        "let r = Sphere { radius: 0.5 }.repeat_x(2.0, 3);\n"
        "let c = Sphere { radius: 1.0 }.slice_y(-0.25, 0.5);\n"
        "let floor = Box { size: vec3(1.0, 1.0, 1.0) };\n"
        "let pu = u.attach(floor, Top);\n"
        "let pm = Sphere { radius: 0.1 }.attach(m, Right);\n"
        "let pr = Sphere { radius: 0.1 }.attach(r, Left);\n"
        "let pc = Sphere { radius: 0.1 }.attach(c, Top);")
    assert _vec(st, "pu")[1] == f32(1.5)          # union bounds y in [-1, 1]
    assert _vec(st, "pm")[0] == f32(3.1)          # mirrored x in [-3, 3]
    assert _vec(st, "pr")[0] == f32(-2.6)         # copies at -2, 0, 2 with radius 0.5
    assert _vec(st, "pc")[1] == f32(0.6)          # sliced to y <= 0.5


# ------------------------------------------------------------------------ imports
def _write(d, name, text):
    p = os.path.join(d, name)
    os.makedirs(os.path.dirname(p), exist_ok=True)
    with open(p, "w") as f:
        f.write(text)
    return p


@pytest.fixture
def library(tmp_path):
    lib = tmp_path / "library"
    _write(str(lib), "materials/gold.ft", "material Gold { model: Metal; color = #ebc757; roughness = 0.2; };\n"
                                          "export { Gold };")
    _write(str(lib), "materials/glass.ft", "material Glass { model: Dielectric; ior = 1.5; };\nexport { Glass };")
    _write(str(lib), "objects/soft_blob.ft",
           "sdf SoftBlob { let r = 1.0; fn bounds() { return vec3(r, r, r); }\n"
           "  fn distance(p) { return length(p) - r; } };\nexport { SoftBlob };")
    return str(lib)


def test_relative_imports_load_once(tmp_path):  # lib.rs:2078-2112
    d = str(tmp_path)
    _write(d, "shared.ft", "let gold = Metal { color: vec3(0.9, 0.7, 0.3), roughness: 0.2 };")
    _write(d, "a.ft", 'import "./shared.ft"; let left = Sphere { material: gold };')
    _write(d, "b.ft", 'import "./shared.ft"; let right = Sphere { material: gold };')
    main = _write(d, "main.ft", 'import "./a.ft"; import "./b.ft"; let scene = left + right;')
    prog = load_program_with_imports(main, [])
    assert sum(1 for st in prog if st[0] == "let" and st[1] == "gold") == 1
    assert "scene" in load_scene_file(main, []).bindings


def test_library_imports_by_path_and_name(tmp_path, library):  # lib.rs:2115-2142
    d = str(tmp_path / "scene")
    main = _write(d, "main.ft", 'import "materials/gold.ft"; import "Glass";\n'
                                "let scene = Sphere { material: Gold {} } + Sphere { material: Glass {} };")
    st = load_scene_file(main, [library])
    assert "Gold" in st.material_defs and "Glass" in st.material_defs
    with pytest.raises(ImportError_):
        load_scene_file(main, [])


def test_namespaced_imports(tmp_path, library):  # lib.rs:2145-2172
    d = str(tmp_path / "scene")
    main = _write(d, "main.ft", 'import "materials/gold.ft" as gold;\nimport "objects/soft_blob.ft" as blob;\n'
                                "let scene = blob.SoftBlob { material: gold.Gold {} };")
    prog = load_program_with_imports(main, [library])
    assert any(st[0] == "def" and st[1].name == "gold.Gold" for st in prog)
    assert any(st[0] == "def" and st[1].name == "blob.SoftBlob" for st in prog)
    st = load_scene_file(main, [library])
    assert "gold.Gold" in st.material_defs and "blob.SoftBlob" in st.sdf_defs
    cs = compile_scene(st)
    assert cs.nodes[cs.root]["kind"] == abi.NODE["CUSTOM"]
    assert cs.materials[cs.nodes[cs.root]["material_id"]]["model"] == abi.MODEL_METAL


def test_same_import_under_two_aliases(tmp_path):  # lib.rs:2175-2199
    d = str(tmp_path)
    _write(d, "shared.ft", "let mat = Metal { color: vec3(0.9, 0.7, 0.3), roughness: 0.2 };")
    main = _write(d, "main.ft", 'import "./shared.ft" as warm;\nimport "./shared.ft" as copy;\n'
                                "let scene = Sphere { material: warm.mat } + Sphere { material: copy.mat };")
    st = load_scene_file(main, [])
    assert {"warm.mat", "copy.mat", "scene"} <= set(st.bindings)


def test_namespaced_functions_and_export_dependencies(tmp_path):  # lib.rs:2202-2289
    d = str(tmp_path)
    _write(d, "helpers.ft", "fn accent() { return #ebc757; }\n"
                            "fn make_gold() { return Metal { color: accent(), roughness: 0.2 }; }\n"
                            "export { make_gold };")
    _write(d, "module.ft", "let private_color = #ebc757;\n"
                           "let unused = 3.0;\n"
                           "let public_mat = Metal { color: private_color, roughness: 0.2 };\n"
                           "export { public_mat };")
    main = _write(d, "main.ft", 'import "./helpers.ft" as palette;\nimport "./module.ft";\n'
                                'import "./module.ft" as lib;\n'
                                "let scene = Sphere { material: palette.make_gold() } + Sphere { material: public_mat }"
                                " + Sphere { material: lib.public_mat };")
    st = load_scene_file(main, [])
    assert {"palette.make_gold", "palette.accent"} <= set(st.function_defs)
    assert {"public_mat", "private_color", "lib.public_mat", "lib.private_color"} <= set(st.bindings)
    assert "unused" not in st.bindings
    compile_scene(st)


def test_import_cycles_are_rejected(tmp_path):  # lib.rs:2292-2300
    d = str(tmp_path)
    _write(d, "a.ft", 'import "./b.ft"; let a = 1.0;')
    _write(d, "b.ft", 'import "./a.ft"; let b = 2.0;')
    with pytest.raises(ImportError_, match="cycle"):
        load_program_with_imports(os.path.join(d, "a.ft"), [])


# --------------------------------------------------------------------------- room
def test_room_compiles_to_its_shown_parts():
    """compile_room (renderer.rs:3359-3520): floor, back wall and right wall by default,
    each a box offset along world axes from the room centre, joined by unions; materials
    interned for every part in order before the parts are built."""
    cs = compile_scene(load_scene_source(
        "var room = Room { width: 6.0, height: 3.0, depth: 5.0, wall_thickness: 0.2,\n"
        "  wall_material: Lambert { color: #ff0000 }, floor_material: Lambert { color: #00ff00 } };\n"
        "room.pos.y = 1.0;\nlet scene = room;"))
    boxes = [n for n in cs.nodes if n["kind"] == abi.NODE["BOX"]]
    unions = [n for n in cs.nodes if n["kind"] == abi.NODE["UNION"]]
    assert len(boxes) == 3 and len(unions) == 2
    centers = sorted(tuple(np.round(cs.object_transforms[b["object_id"], :3], 5)) for b in boxes)
    assert centers == sorted([(0.0, -0.4, 0.0), (0.0, 1.0, -2.4), (2.9, 1.0, 0.0)])
    halves = sorted(tuple(np.round(b["param"][:3], 5)) for b in boxes)
    assert halves == sorted([(3.0, 0.1, 2.5), (3.0, 1.5, 0.1), (0.1, 1.5, 2.5)])
    floor = [b for b in boxes if abs(cs.object_transforms[b["object_id"], 1] + 0.4) < 1e-6][0]
    assert np.allclose(cs.materials[floor["material_id"]]["color"], [0, 1, 0])
    st = load_scene_source("let room = Room { width: 6.0, height: 3.0, depth: 5.0 };\n"
                           "let s = Sphere { radius: 0.5 }.attach(room, Top);")
    assert _vec(st, "s")[1] == f32(2.0)


# ------------------------------------------------------------------------ skeletons
_RIG = ("skeleton Rig {\n  joint a = vec3(0.0, 1.6, 0.0);\n  joint b = vec3(-0.3, 1.0, 0.0);\n"
        "  bone arm = a, b;\n};\n")


def test_skeleton_instances_and_part_proxies():  # lib.rs:2589-2635
    st = load_scene_source(_RIG + "let rig = Rig {};\nlet j = rig.b;\nlet arm = rig.arm;")
    rig = st.bindings["rig"].value
    assert {"anchors", "__skeleton_joints", "__skeleton_bones", "__bounds"} <= set(rig.fields)
    assert _vec(st, "rig") == [f32(-0.15), f32(1.3), f32(0.0)]      # bounds centre
    j = st.bindings["j"].value
    assert j.type_name == "SkeletonJoint" and "__bounds" in j.fields
    assert _vec(st, "j") == [f32(-0.3), f32(1.0), f32(0.0)]
    arm = st.bindings["arm"].value
    assert arm.type_name == "SkeletonBone" and {"Start", "End"} <= set(arm.fields["anchors"].fields)


def test_bind_aligns_box_long_axis_to_bone():  # lib.rs:2638-2755
    from forgedthoughts_amd.forge.layout import rotate_xyz
    st = load_scene_source(_RIG + "let rig = Rig {};\nlet part = Box { size: vec3(0.1, 0.1, 0.5) }.bind(rig.arm);")
    pos, rot, size = (np.array(_vec(st, "part", k), dtype=np.float64) for k in ("pos", "rot", "size"))
    half = size[2] * 0.5
    ends = [pos + np.array(rotate_xyz([f32(0), f32(0), f32(s * half)], [f32(r) for r in rot]), dtype=np.float64)
            for s in (-1, 1)]
    a, b = np.array([0.0, 1.6, 0.0]), np.array([-0.3, 1.0, 0.0])
    best = min(np.linalg.norm(ends[0] - a) + np.linalg.norm(ends[1] - b),
               np.linalg.norm(ends[0] - b) + np.linalg.norm(ends[1] - a))
    assert best < 0.08
    assert abs(size[2] - 1.03 * np.hypot(0.3, 0.6)) < 1e-5      # stretched to 1.03 x the bone


def test_ik_keeps_two_bone_chain_lengths():  # lib.rs:2758-2814
    st = load_scene_source("skeleton Rig {\n  joint shoulder = vec3(0.0, 1.6, 0.0);\n"
                           "  joint elbow = vec3(-0.2, 1.0, 0.0);\n  joint hand = vec3(-0.4, 0.5, 0.0);\n"
                           "  chain arm = shoulder, elbow, hand;\n};\n"
                           "let rig = Rig { ik: { arm: vec3(-0.6, 0.1, 0.2) } };")
    joints = st.bindings["rig"].value.fields["__skeleton_joints"].fields
    j = {k: np.array([float(v.fields[c]) for c in "xyz"]) for k, v in joints.items()}
    assert abs(np.linalg.norm(j["shoulder"] - j["elbow"]) - 0.6324555) < 0.02
    assert abs(np.linalg.norm(j["elbow"] - j["hand"]) - 0.53851646) < 0.02
    assert np.linalg.norm(j["hand"] - j["elbow"] - (np.array([-0.2, -0.5, 0.0]))) > 0.05   # the chain moved


def test_native_sdf_helpers_fold_and_lower_alike():
    """box_shell_sdf / cylinder_*_sdf / hole_line_*_sdf (eval.rs:1943-2009): on numbers they
    fold to the closed forms; inside a custom distance function they lower to the expression
    DAG, which the oracle then renders."""
    from oracle import ftref
    st = load_scene_source("let a = box_shell_sdf(vec3(0.0, 0.0, 2.0), vec3(1.0, 1.0, 1.0), 0.1, 0.05);\n"
                           "let b = cylinder_y_sdf(vec3(3.0, 0.0, 0.0), 1.0, 2.0);\n"
                           "let c = cylinder_x_sdf(vec3(0.0, 0.0, 0.0), 0.5, 1.0);\n"
                           "let d = hole_line_z_sdf(vec3(0.0, 0.0, 0.3), 0.1, 1.0, 1.0, 3.0);")
    v = {k: float(st.bindings[k].value) for k in "abcd"}
    assert abs(v["a"] - 1.0) < 1e-6          # outside the outer box face at z = 1
    assert abs(v["b"] - 2.0) < 1e-6          # radial distance 3 - 1
    assert abs(v["c"] + 0.5) < 1e-6          # inside: max(-0.5, -1) = -0.5
    assert abs(v["d"] + 0.1) < 1e-6          # on the axis of z-holes shifted along z: -radius
    cs = compile_scene(load_scene_source(
        "sdf Part { fn bounds() { return vec3(1.5, 1.5, 1.5); }\n"
        "  fn distance(p) { return min(box_shell_sdf(p, vec3(1.0, 0.8, 0.6), 0.1, 0.05),\n"
        "                            hole_line_x_sdf(p - vec3(0.0, 0.9, 0.0), 0.08, 0.6, 0.3, 4.0)); } };\n"
        "let camera = Camera { origin: vec3(2.0, 1.5, 4.0), target: vec3(0.0, 0.0, 0.0), fov_y: 40.0 };\n"
        "let scene = Part {};"))
    from forgedthoughts_amd.options import RenderOptions
    _, t = ftref.render_depth(cs, 1, RenderOptions(32, 24, 256), 1)
    assert (t >= 0).mean() > 0.1


def test_layout_against_semantic_part_proxies():  # lib.rs:1534-1558
    st = load_scene_source("var table = Table { width: 1.6, depth: 0.8, height: 0.75, top_thickness: 0.1 };\n"
                           "var vase = Sphere { radius: 0.2 }.attach(table.top, Top);")
    assert _vec(st, "vase")[1] == f32(0.575)


def test_semantic_assets_compile_to_their_parts():
    """compile_lowered_library_object (renderer.rs:2413-2770): a Table is a top box and four
    leg cylinders, a Cupboard a shell (box minus cavity) and a door, a Lamp three cylinders
    and a bulb sphere plus a sphere light at the bulb."""
    defs = ("sdf Table { fn bounds() { return vec3(1.0); } fn distance(p) { return length(p) - 1.0; } };\n"
            "sdf Cupboard { fn bounds() { return vec3(1.0); } fn distance(p) { return length(p) - 1.0; } };\n"
            "sdf Lamp { fn bounds() { return vec3(1.0); } fn distance(p) { return length(p) - 1.0; } };\n")
    kinds = lambda cs: sorted(abi.NODE_KINDS[int(n["kind"])] for n in cs.nodes)  # noqa: E731
    t = compile_scene(load_scene_source(defs + "let scene = Table { width: 1.6, height: 0.75 };"))
    assert kinds(t) == sorted(["BOX"] + ["CYLINDER"] * 4 + ["UNION"] * 4)
    top = [n for n in t.nodes if n["kind"] == abi.NODE["BOX"]][0]
    assert abs(float(t.object_transforms[top["object_id"], 1]) - (0.375 - 0.04)) < 1e-6
    c = compile_scene(load_scene_source(defs + "let scene = Cupboard { };"))
    assert kinds(c) == sorted(["BOX"] * 3 + ["SUBTRACT", "UNION"])
    n_lights = len(compile_scene(load_scene_source(defs + "let k = PointLight { position: vec3(0.0, 3.0, 0.0) };\n"
                                                          "let scene = Sphere { radius: 1.0 };")).lights)
    lamp = compile_scene(load_scene_source(defs + "let k = PointLight { position: vec3(0.0, 3.0, 0.0) };\n"
                                                  "var lamp = Lamp { height: 0.72 }; lamp.pos.x = 1.0;\nlet scene = lamp;"))
    assert kinds(lamp) == sorted(["CYLINDER"] * 3 + ["SPHERE"] + ["UNION"] * 3)
    assert len(lamp.lights) == n_lights + 1
    bulb = lamp.lights[-1]
    assert bulb["kind"] == abi.LIGHT_SPHERE and abs(float(bulb["position"][0]) - 1.0) < 1e-6
    assert np.allclose(bulb["radiance"], [18.0, 17.2, 15.6])
