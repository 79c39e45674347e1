"""The reference's example scenes (examples/*.ft) through the GPU backend.

The scenes are compiled here once, by tests/golden/make_examples.py, into IR fixtures
(tests/golden/examples/*.npz), so these tests need no checkout of the reference.  GPU
tests render every example's depth preview and Whitted image on the GPU and with the
oracle at a reduced size and compare them; CPU tests check the fixture format and, where
a checkout is present, that the fixtures match what the front end compiles today."""
import glob
import os

import numpy as np
import pytest

from forgedthoughts_amd.options import AccelMode, RaySettings, RenderOptions
from forgedthoughts_amd.scene import CompiledScene, compile_scene
from helpers import ROOT, compile_src
from oracle import ftref

EXAMPLES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "examples", "*.npz")))
NAMES = [os.path.basename(p)[:-4] for p in EXAMPLES]
REFERENCE_EXAMPLES = "/root/reference/examples"
REFERENCE_LIBRARY = "/root/reference/crates/forgedthoughts/library"


def _load(name):
    scene, extra = CompiledScene.load(os.path.join(ROOT, "tests", "golden", "examples", name + ".npz"))
    o = extra["options"]
    opts = RenderOptions(64, 48, int(o[0]), float(o[1]), float(o[2]), float(o[3]), float(o[4]), float(o[5]))
    return scene, opts, AccelMode(int(extra["accel"]))


def test_example_fixtures_present():
    assert len(EXAMPLES) >= 45
    for n in ("boolean_union_round", "boolean_tongue", "stone", "room", "robot_on_stone", "imports", "ft_material_wax"):
        assert n in NAMES


def test_ir_file_round_trip(tmp_path):
    """CompiledScene.save / load keep every IR array and the camera: the oracle renders the
    loaded scene to the same bytes."""
    scene = compile_src("let camera = Camera { origin: vec3(0.0, 1.0, 5.0), target: vec3(0.0, 0.0, 0.0), fov_y: 40.0 };\n"
                        "material M { model: Lambert; fn color(ctx) { return ctx.normal * 0.5 + vec3(0.5); } };\n"
                        "let scene = Sphere { radius: 1.0, material: M {} }.union_round(Box { size: vec3(1.5) }, 0.2);")
    p = str(tmp_path / "s.npz")
    scene.save(p, options=np.arange(3, dtype=np.float32))
    back, extra = CompiledScene.load(p)
    for k in CompiledScene._ARRAYS:
        assert getattr(scene, k).tobytes() == getattr(back, k).tobytes(), k
    assert np.array_equal(extra["options"], np.arange(3, dtype=np.float32))
    opts = RenderOptions(24, 18, 128)
    a, _ = ftref.render_whitted(scene, AccelMode.Bvh, opts, RaySettings(2, 24, 1))
    b, _ = ftref.render_whitted(back, AccelMode.Bvh, opts, RaySettings(2, 24, 1))
    assert np.array_equal(a, b) and a.max() > 0


@pytest.mark.skipif(not os.path.isdir(REFERENCE_EXAMPLES), reason="no reference checkout here")
@pytest.mark.parametrize("name", NAMES)
def test_example_fixture_is_current(name):
    """The fixture equals what the front end compiles from the example today."""
    from forgedthoughts_amd.forge.imports import load_scene_file
    fresh = compile_scene(load_scene_file(os.path.join(REFERENCE_EXAMPLES, name + ".ft"), [REFERENCE_LIBRARY]))
    saved, _ = CompiledScene.load(os.path.join(ROOT, "tests", "golden", "examples", name + ".npz"))
    for k in CompiledScene._ARRAYS:
        assert getattr(fresh, k).tobytes() == getattr(saved, k).tobytes(), f"{name}: {k} differs"


def _whitted_refused(scene) -> bool:
    """A material with per-hit hooks but no hook table entry (bump / normal / eval / medium
    hooks, which the GPU does not evaluate): Whitted renders of the scene fail loudly."""
    hooked = set(int(m) for m in scene.material_hooks["material"])
    return any(int(m["dynamic_material_id"]) >= 0 and i not in hooked for i, m in enumerate(scene.materials))


@pytest.mark.gpu
@pytest.mark.parametrize("emit", ["auto", "0"], ids=["default", "interp"])
@pytest.mark.parametrize("name", NAMES)
def test_example_gpu_parity(name, emit, monkeypatch):
    """Depth preview and Whitted (3 bounces, 1 spp) of one example, GPU against the oracle,
    at 64x48 with the example's own march settings, on the default path (scene-specialised
    code where the scene allows it) and on the interpreter; Whitted of an example with hooks
    the GPU does not evaluate must raise UnsupportedObjectType."""
    monkeypatch.setenv("FT_EMIT", emit)
    from forgedthoughts_amd.render import GpuScene
    from forgedthoughts_amd.scene import RenderError
    from test_gpu_parity import depth_parity, whitted_parity
    scene, opts, accel = _load(name)
    # the refine step bisects the last march step six times, so when GPU and CPU rounding
    # lead to march steps of different lengths the hit t can differ by a fraction of the
    # step; the examples' own epsilon bounds that for their long custom distance functions
    depth_parity(scene, accel, opts, 1, name, t_tol=max(1.0e-4, opts.epsilon))
    if _whitted_refused(scene):
        with GpuScene(scene, accel) as gs:
            with pytest.raises(RenderError) as e:
                gs.render_whitted(opts, RaySettings(3, 64, 1))
            assert e.value.kind == "UnsupportedObjectType"
        return
    # at 64x48 one pixel is 0.03% of the image: a few grazing secondary rays (reflection or
    # refraction at a silhouette) that take the other branch on GPU and CPU rounding sink a
    # 99.9% bar, so these small renders take 99.5% within 1/255 and the reference's own mean
    # bar for its render regressions (mean |diff| <= 0.35, render_regression.rs:138-145)
    rgb, _ = whitted_parity(scene, accel, opts, RaySettings(3, 64, 1), name, min_frac=0.995)
    ref, _ = ftref.render_whitted(scene, accel, opts, RaySettings(3, 64, 1))
    mean = float(np.abs(rgb.astype(int) - ref.astype(int)).mean())
    assert mean <= 0.35, f"{name}: mean byte difference {mean:.4f}"
