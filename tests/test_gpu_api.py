"""The reference's entry points on the GPU backend (renderer.rs:1139-1277), through the
Python mirror of the API and the C ABI's progressive renders: images against the oracle,
the progress callback on the caller's thread per tile in the reference's order, and a
callback error stopping the render."""
import os
import threading

import numpy as np
import pytest

from forgedthoughts_amd import render
from forgedthoughts_amd.forge.evaluator import load_scene_source
from forgedthoughts_amd.options import AccelMode, RaySettings, RenderOptions
from forgedthoughts_amd.scene import RenderError, compile_scene
from oracle import ftref

from helpers import REGRESSION, REGRESSION_CASES, diff_metrics, read_png

pytestmark = pytest.mark.gpu


def _state(name):
    with open(os.path.join(REGRESSION, name + ".ft")) as f:
        return load_scene_source(f.read())


def _tile_order(w, h, ts):
    return [(x, y) for y in range(0, h, ts) for x in range(0, w, ts)]


def test_render_depth_png_with_accel_matches_reference_png(tmp_path):
    name, _, opts, _ = REGRESSION_CASES[0]
    out = str(tmp_path / "depth.png")
    render.render_depth_png_with_accel(_state(name), out, opts, AccelMode.Bvh)
    got = read_png(out)
    assert diff_metrics(got, read_png(os.path.join(REGRESSION, name + ".png"))) == (0, 0.0)


def test_preview_progressive_calls_back_per_tile():
    name, _, opts, _ = REGRESSION_CASES[3]
    state = _state(name)
    seen, threads = [], set()

    def on_tile(progress, image):
        seen.append((progress["tiles_done"], progress["tiles_total"], progress["tile"][:2], image))
        threads.add(threading.get_ident())
    rgb = render.render_preview_progressive_with_accel(state, opts, AccelMode.Bvh, 40, 1, on_tile)
    ref, _ = ftref.render_depth(compile_scene(state), AccelMode.Bvh, opts, 1)
    assert diff_metrics(rgb, ref) == (0, 0.0)
    order = _tile_order(opts.width, opts.height, 40)
    assert [s[0] for s in seen] == list(range(1, len(order) + 1))
    assert all(s[1] == len(order) for s in seen)
    assert [s[2] for s in seen] == order
    assert threads == {threading.get_ident()}
    np.testing.assert_array_equal(seen[-1][3], rgb)
    # tiles not yet finished are black in earlier images
    x, y = order[-1]
    assert not seen[0][3][y:y + 40, x:x + 40].any() or not ref[y:y + 40, x:x + 40].any()


@pytest.mark.parametrize("spp,tile", [(2, 32), (4, 30)])
def test_ray_progressive_matches_the_whole_frame_render(spp, tile):
    """Bands of rows against the whole frame: at 2 spp the default Whitted shape falls back to
    8x8 pixel tiles, at 4 spp its sample-major waves cover 4x4-pixel blocks, which 30-row
    bands cut."""
    name, _, opts, settings = REGRESSION_CASES[2]
    state = _state(name)
    s = RaySettings(settings.max_depth, tile, spp)
    count = []
    rgb = render.render_ray_progressive_with_accel(state, opts, AccelMode.Bvh, s, lambda p, im: count.append(p))
    with render.GpuScene(compile_scene(state), AccelMode.Bvh) as gs:
        whole, _ = gs.render_whitted(opts, s)
    np.testing.assert_array_equal(rgb, whole)
    assert len(count) == len(_tile_order(opts.width, opts.height, tile))
    ref, _ = ftref.render_whitted(compile_scene(state), AccelMode.Bvh, opts, s)
    d = np.abs(rgb.astype(int) - ref.astype(int)).max(axis=2)
    assert float((d <= 1).mean()) >= 0.999


def test_callback_error_stops_the_render():
    """An Err from on_tile ends the render with that error (renderer.rs:3695-3716)."""
    name, _, opts, settings = REGRESSION_CASES[1]
    state = _state(name)
    calls = []

    def on_tile(progress, image):
        calls.append(progress["tiles_done"])
        if progress["tiles_done"] == 3:
            raise RenderError("Image", "stop here")
    with pytest.raises(RenderError) as e:
        render.render_ray_progressive_with_accel(state, opts, AccelMode.Bvh, RaySettings(2, 16, 1), on_tile)
    assert e.value.kind == "Image" and "stop here" in str(e.value)
    assert calls == [1, 2, 3]
    with pytest.raises(RenderError) as e:
        render.render_preview_progressive_with_accel(state, opts, AccelMode.Naive, 16, 1, on_tile)
    assert calls[-1] == 3


def test_ftc_depth_trace_and_bench(tmp_path):
    """`ftc depth` and `ftc` (trace) write the images the oracle renders with the options
    the CLI merges (scene settings under --width / --height, M:792-829); `ftc bench` times
    the three accel modes."""
    from forgedthoughts_amd import ftc
    from forgedthoughts_amd.forge.imports import load_scene_file
    from forgedthoughts_amd.scene import extract_scene_render_settings
    name, _, _, _ = REGRESSION_CASES[0]
    scene = os.path.join(REGRESSION, name + ".ft")
    state = load_scene_file(scene)
    opts = ftc.merged_render_options(extract_scene_render_settings(state), 80, 64)
    out = str(tmp_path / "d.png")
    assert ftc.main(["depth", "--scene", scene, "--output", out, "--width", "80", "--height", "64"]) == 0
    ref, _ = ftref.render_depth(compile_scene(state), AccelMode.Bvh, opts, 1)
    assert diff_metrics(read_png(out), ref) == (0, 0.0)
    name2 = REGRESSION_CASES[1][0]
    scene2 = os.path.join(REGRESSION, name2 + ".ft")
    state2 = load_scene_file(scene2)
    opts2 = ftc.merged_render_options(extract_scene_render_settings(state2), 80, 64)
    out2 = str(tmp_path / "t.png")
    assert ftc.main(["--scene", scene2, "--output", out2, "--width", "80", "--height", "64", "--depth", "4"]) == 0
    ref2, _ = ftref.render_whitted(compile_scene(state2), AccelMode.Bvh, opts2, RaySettings(4, 64, 1))
    d = np.abs(read_png(out2).astype(int) - ref2.astype(int)).max(axis=2)
    assert float((d <= 1).mean()) >= 0.999
    res = ftc.run_bench(ftc.parser().parse_args(["bench", "--scene", scene, "--iterations", "2", "--warmup", "1"]))
    assert res["exit"] == 0
    assert [r["accel"] for r in res["rows"]] == ["Naive", "Bvh", "Bricks"]
    assert all(r["best_ms"] <= r["avg_ms"] + 1 and r["mpix_per_sec"] > 0 for r in res["rows"])
