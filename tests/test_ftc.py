"""The ftc driver (forgedthoughts_amd/ftc.py) against the reference CLI's contract
(crates/ftc/src/main.rs): exit codes, scene resolution, option merging, output naming and
the watch loop.  GPU renders through it are in test_gpu_api.py."""
import os

import pytest

from forgedthoughts_amd import ftc
from forgedthoughts_amd.options import AccelMode

from helpers import REGRESSION

SCENE = os.path.join(REGRESSION, "preview_lambert.ft")


def test_missing_scene_is_exit_2(monkeypatch):
    monkeypatch.delenv("FORGEDTHOUGHTS_SCENE", raising=False)
    assert ftc.main(["depth"]) == 2
    assert ftc.main([]) == 2
    assert ftc.main(["bench"]) == 2
    assert ftc.main(["check"]) == 2


def test_zero_iterations_is_exit_2():
    assert ftc.main(["bench", "--scene", SCENE, "--iterations", "0"]) == 2


def test_bad_scene_is_exit_3(tmp_path):
    bad = tmp_path / "bad.ft"
    bad.write_text("let scene = Sphere { radius: ; };")
    assert ftc.main(["check", "--scene", str(bad)]) == 3
    assert ftc.main(["depth", "--scene", str(bad)]) == 3
    assert ftc.main(["--scene", str(bad)]) == 3
    assert ftc.main(["bench", "--scene", str(bad)]) == 3


def test_check_uses_the_environment_scene(monkeypatch):
    monkeypatch.setenv("FORGEDTHOUGHTS_SCENE", SCENE)
    assert ftc.main(["check"]) == 0


def test_output_path_and_option_merging():
    assert ftc.default_output_path("a/b/scene.ft") == "a/b/scene.png"
    assert ftc.default_output_path("a/b/scene") == "a/b/scene.png"
    from forgedthoughts_amd.scene import SceneRenderSettings
    s = SceneRenderSettings(width=320, height=200, max_steps=99, step_scale=3.0, fov_y_degrees=30.0)
    o = ftc.merged_render_options(s, None, 100)
    assert (o.width, o.height, o.max_steps, o.step_scale, o.fov_y_degrees) == (320, 100, 99, 1.0, 30.0)
    assert o.max_dist == 40.0 and o.epsilon == pytest.approx(0.001)
    s.accel = 2
    assert ftc._accel(None, s) == AccelMode.Bricks and ftc._accel("naive", s) == AccelMode.Naive
    s.accel = None
    assert ftc._accel(None, s) == AccelMode.Bvh


def test_watch_reruns_on_change(tmp_path):
    f = tmp_path / "s.ft"
    f.write_text("let scene = Sphere { radius: 1.0 };")
    runs = []

    def sleep(_):
        if len(runs) == 1:
            f.write_text("let scene = Sphere { radius: 2.0 };  ")
    assert ftc.run_with_watch(str(f), True, lambda p: runs.append(p) or 0, sleep=sleep, max_polls=3) == 0
    assert len(runs) == 2


def test_render_without_backend_is_exit_4(tmp_path, monkeypatch):
    """With no GPU (or no built library) the render fails loudly: exit 4, no CPU fallback."""
    import forgedthoughts_amd.render as R
    try:
        if R.device_count() > 0:
            pytest.skip("a GPU is present")
    except Exception:
        pass
    assert ftc.main(["depth", "--scene", SCENE, "--output", str(tmp_path / "o.png")]) == 4
