"""Forge front end: parsing, evaluation and scene compilation (host logic, no GPU)."""
import os

import numpy as np
import pytest

from forgedthoughts_amd import abi
from forgedthoughts_amd.forge.evaluator import ForgeEvalError, load_scene_source
from forgedthoughts_amd.forge.syntax import ForgeSyntaxError, parse_program
from forgedthoughts_amd.scene import RenderError, compile_scene, extract_scene_render_settings

from helpers import REGRESSION, REGRESSION_CASES, compile_ft, compile_src


@pytest.mark.parametrize("name", [c[0] for c in REGRESSION_CASES])
def test_regression_scenes_compile(name):
    cs = compile_ft(os.path.join(REGRESSION, name + ".ft"))
    assert cs.nodes[cs.root]["kind"] == abi.NODE["UNION"]
    assert cs.env_fn >= 0
    assert cs.camera["present"]


def test_hex_colors_are_byte_over_255():
    st = load_scene_source("let c = #ff8000; let d = #0f0;")
    c = st.bindings["c"].value.fields
    assert float(c["x"]) == np.float32(1.0) and float(c["y"]) == np.float32(128 / 255.0) and float(c["z"]) == 0.0
    d = st.bindings["d"].value.fields
    assert float(d["y"]) == 1.0


def test_precedence_and_vector_math():
    st = load_scene_source("let a = 1 + 2 * 3 - 4 / 2; let v = vec3(1, 2, 3) * 2 + 1; let n = -v.x;")
    assert float(st.bindings["a"].value) == 5.0
    v = st.bindings["v"].value.fields
    assert [float(v[k]) for k in "xyz"] == [3.0, 5.0, 7.0]
    assert float(st.bindings["n"].value) == -3.0


def test_parse_errors():
    with pytest.raises(ForgeSyntaxError):
        parse_program("let a = 1")
    with pytest.raises(ForgeSyntaxError):
        parse_program("let c = #12345;")


def test_immutable_assignment_rejected():
    with pytest.raises(ForgeEvalError):
        load_scene_source("let a = Sphere { radius: 1.0 }; a.pos.x = 1.0;")


def test_missing_scene_root():
    with pytest.raises(RenderError) as e:
        compile_scene(load_scene_source("let a = 1.0;"))
    assert e.value.kind == "MissingSceneRoot"


def test_unsupported_object_type():
    with pytest.raises(RenderError) as e:
        compile_scene(load_scene_source("let scene = Teapot { };"))
    assert e.value.kind == "UnsupportedObjectType"


def test_operators_compile_to_nodes():
    src = """
    let a = Sphere { radius: 1.0 };
    let b = Box { size: vec3(1.0, 2.0, 3.0), round: 0.1 };
    let scene = a.union_round(b, 0.2) - Torus {} & Cylinder { height: 3.0 };
    """
    cs = compile_scene(load_scene_source(src))
    kinds = [abi.NODE_KINDS[k] for k in cs.nodes["kind"]]
    assert kinds == ["SPHERE", "BOX", "UNION_ROUND", "TORUS", "SUBTRACT", "CYLINDER", "INTERSECT"]
    box = cs.nodes[1]
    assert list(box["param"][:5]) == pytest.approx([0.5, 1.0, 1.5, 0.1, 0.0])
    assert cs.nodes[5]["param"][1] == pytest.approx(1.5)  # half height


def test_render_settings_extraction():
    st = load_scene_source("let render = RenderSettings { width: 320, height: 200, max_steps: 64, "
                           "epsilon: 0.002, accel: Bricks {} };")
    rs = extract_scene_render_settings(st)
    assert (rs.width, rs.height, rs.max_steps, rs.accel) == (320, 200, 64, abi.ACCEL_BRICKS)
    assert rs.epsilon == pytest.approx(0.002)


def test_default_lights_added():
    cs = compile_scene(load_scene_source("let scene = Sphere {};"))
    kinds = list(cs.lights["kind"])
    assert kinds == [abi.LIGHT_POINT, abi.LIGHT_ENV]
    assert list(cs.lights[0]["radiance"]) == [6.0, 6.0, 6.0]


def test_repeat_and_mirror():
    cs = compile_scene(load_scene_source("let b = Sphere { radius: 0.2 }; let scene = b.repeat_x(1.0, 3).mirror_y();"))
    prims = cs.nodes[cs.nodes["kind"] == abi.NODE["SPHERE"]]
    xs = sorted(set(float(x) for x in prims["center"][:, 0]))
    assert xs == pytest.approx([-1.0, 0.0, 1.0])


def test_per_hit_material_hooks_are_recorded():
    """Only material definitions with per-hit hook functions are marked dynamic
    (resolve_leaf_material renderer.rs:7164-7536); helper functions are not hooks."""
    s = compile_ft(os.path.join(REGRESSION, "preview_material_hook.ft"))
    assert s.dynamic_hooks == ["StripeLambert.color"]
    assert sorted(s.materials["dynamic_material_id"].tolist()) == [-1, 0]
    s = compile_src("material Plain { model: Lambert; let c = #ff0000; color = c; };\n"
                    "let scene = Sphere { radius: 1.0, material: Plain {} };")
    assert s.dynamic_hooks == [] and (s.materials["dynamic_material_id"] == -1).all()


def test_standard_definition_coat_properties():
    """Standard / Layered definitions take their specular colour, ior and clearcoat from the
    coat properties at every hit (renderer.rs:7386-7470): defaults white, 1.33, 0 x 1, 0.08."""
    s = compile_src("material P { model: Standard; color = #ff0000; ior = 1.8; specular_color = #00ff00; };\n"
                    "let scene = Sphere { radius: 1.0, material: P {} };")
    m = s.materials[int(s.nodes["material_id"][0])]
    assert m["ior"] == np.float32(1.33) and tuple(m["specular_color"]) == (1.0, 1.0, 1.0)
    assert m["clearcoat"] == 0.0 and m["clearcoat_roughness"] == np.float32(0.08)
    assert tuple(m["color"]) == (1.0, 0.0, 0.0)
    s = compile_src("material C { model: Layered; coat_weight = 0.5; coat_ior = 1.6; coat_roughness = 0.2;\n"
                    "  coat_color = vec3(0.5, 0.6, 0.7); };\nlet scene = Sphere { radius: 1.0, material: C {} };")
    m = s.materials[int(s.nodes["material_id"][0])]
    assert m["ior"] == np.float32(1.6) and m["clearcoat"] == np.float32(0.5)
    assert m["clearcoat_roughness"] == np.float32(0.2)
    assert np.allclose(m["specular_color"], (0.5, 0.6, 0.7))
