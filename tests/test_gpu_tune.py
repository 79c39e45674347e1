"""Launch shapes and the autotuner (ft_scene_set_launch_config, ft_scene_autotune): every
shape renders the oracle's image, and the tuner's choice is a valid shape that it reuses."""
import numpy as np
import pytest

from forgedthoughts_amd import abi, synthetic
from forgedthoughts_amd.options import RaySettings, RenderOptions
from forgedthoughts_amd.render import GpuScene
from oracle import ftref

from helpers import compile_src

pytestmark = pytest.mark.gpu

SHAPES = [(4, 8, 4, 1, 0), (1, 16, 4, 1, 0), (2, 32, 4, 1, 0), (4, 8, 2, 1, 0), (4, 16, 3, 2, 0),
          (1, 8, 4, 4, 0), (4, 8, 4, 1, 1), (2, 16, 2, 1, 1)]
# sample-major Whitted shapes (tile width 0) and the sample counts they apply to
SAMPLE_MAJOR = [(1, 0, 4, 1, 0), (4, 0, 2, 2, 1), (2, 0, 3, 1, 0), (1, 0, 5, 1, 0)]


@pytest.fixture(scope="module")
def c2_small():
    w = synthetic.whitted_1080p(width=160, height=96, spp=2)
    scene = compile_src(w.source)
    rgb_c, _ = ftref.render_whitted(scene, w.accel, w.options, w.settings)
    _, t_c = ftref.render_depth(scene, w.accel, w.options, 1)
    return w, scene, rgb_c, t_c


@pytest.mark.parametrize("shape", SHAPES, ids=[f"bw{a}_t{b}_o{c}_u{d}_i{e}" for a, b, c, d, e in SHAPES])
def test_every_launch_shape_renders_the_oracle_image(c2_small, shape):
    w, scene, rgb_c, t_c = c2_small
    print(f"[shape] {shape}", flush=True)
    with GpuScene(scene, w.accel) as gs:
        gs.set_launch_config(abi.KIND_WHITTED, *shape)
        gs.set_launch_config(abi.KIND_DEPTH, *shape)
        rgb, _ = gs.render_whitted(w.options, w.settings)
        _, t = gs.render_depth(w.options, 1)
        cfg = gs.launch_config(abi.KIND_WHITTED)
    assert (cfg.block_waves, cfg.tile_width, cfg.waves_per_eu, cfg.march_unroll, cfg.shade_inline) == shape
    d = np.abs(rgb.astype(int) - rgb_c.astype(int)).max(axis=2)
    assert float((d <= 1).mean()) >= 0.999, f"{shape}: {(d <= 1).mean():.5f} within 1/255"
    hit, hit_c = t >= 0, t_c >= 0
    assert int((hit != hit_c).sum()) <= max(1, hit.size // 10000)
    both = hit & hit_c
    assert float(np.abs(t[both] - t_c[both]).max()) <= 1e-4


def test_default_shapes_one_wave_blocks_and_sample_major_whitted():
    scene = compile_src("let scene = Sphere { radius: 1.0 };")
    with GpuScene(scene, abi.ACCEL_BVH) as gs:
        gs.set_launch_config(abi.KIND_WHITTED, 4, 16)
        gs.set_launch_config(abi.KIND_DEPTH, 4, 16)
        gs.set_launch_config(abi.KIND_WHITTED)  # the Python defaults restore the library's
        gs.set_launch_config(abi.KIND_DEPTH)
        w = gs.launch_config(abi.KIND_WHITTED)
        d = gs.launch_config(abi.KIND_DEPTH)
    assert (w.block_waves, w.tile_width, w.waves_per_eu, w.march_unroll, w.shade_inline) == (1, 0, 4, 1, 0)
    assert (d.block_waves, d.tile_width, d.waves_per_eu, d.march_unroll, d.shade_inline) == (1, 8, 4, 1, 0)


@pytest.mark.parametrize("spp", [4, 16, 64, 9])
@pytest.mark.parametrize("shape", SAMPLE_MAJOR, ids=[f"bw{a}_t{b}_o{c}_u{d}_i{e}" for a, b, c, d, e in SAMPLE_MAJOR])
def test_sample_major_waves_render_the_pixel_tile_image(shape, spp):
    """A sample-major wave traces the rays an 8x8-tile wave traces, so the image is the same byte
    for byte (spp 9 has no sample-major form and falls back to 8x8 tiles); ragged 61x37 image
    edges leave partly filled waves."""
    w = synthetic.whitted_1080p(width=61, height=37, spp=spp)
    scene = compile_src(w.source)
    print(f"[sample-major] {shape} spp {spp}", flush=True)
    with GpuScene(scene, w.accel) as gs:
        gs.set_launch_config(abi.KIND_WHITTED, 4, 8, shape[2], shape[3], shape[4])
        rgb_t, _ = gs.render_whitted(w.options, w.settings)
        gs.set_launch_config(abi.KIND_WHITTED, *shape)
        rgb_s, _ = gs.render_whitted(w.options, w.settings)
    assert np.array_equal(rgb_s, rgb_t)
    if spp == 4:
        rgb_c, _ = ftref.render_whitted(scene, w.accel, w.options, w.settings)
        d = np.abs(rgb_s.astype(int) - rgb_c.astype(int)).max(axis=2)
        assert float((d <= 1).mean()) >= 0.999


def test_autotune_chooses_a_shape_and_caches_it(c2_small):
    w, scene, rgb_c, _ = c2_small
    with GpuScene(scene, w.accel) as gs:
        rep = gs.autotune(abi.KIND_WHITTED, w.options, w.settings, abi.TUNE_NO_CACHE)
        assert rep.candidates >= 12 and not rep.cached
        assert rep.best_ms > 0 and rep.default_ms > 0 and rep.best_ms <= rep.default_ms * 1.05
        cfg = gs.launch_config(abi.KIND_WHITTED)
        assert (cfg.block_waves, cfg.tile_width, cfg.waves_per_eu, cfg.march_unroll, cfg.shade_inline) == \
            (rep.best.block_waves, rep.best.tile_width, rep.best.waves_per_eu, rep.best.march_unroll,
             rep.best.shade_inline)
        rgb, _ = gs.render_whitted(w.options, w.settings)
    d = np.abs(rgb.astype(int) - rgb_c.astype(int)).max(axis=2)
    assert float((d <= 1).mean()) >= 0.999
    with GpuScene(scene, w.accel) as gs:
        again = gs.autotune(abi.KIND_WHITTED, w.options, w.settings)
        assert again.cached
        assert (again.best.block_waves, again.best.tile_width, again.best.waves_per_eu, again.best.march_unroll,
                again.best.shade_inline) == (rep.best.block_waves, rep.best.tile_width, rep.best.waves_per_eu,
                                             rep.best.march_unroll, rep.best.shade_inline)


def test_invalid_launch_shape_is_refused():
    from forgedthoughts_amd.scene import RenderError
    scene = compile_src("let scene = Sphere { radius: 1.0 };")
    with GpuScene(scene) as gs:
        with pytest.raises(RenderError):
            gs.set_launch_config(abi.KIND_DEPTH, 3, 8, 4, 1, 0)
        with pytest.raises(RenderError):
            gs.set_launch_config(abi.KIND_DEPTH, 4, 12, 4, 1, 0)
        with pytest.raises(RenderError):
            gs.set_launch_config(abi.KIND_DEPTH, 1, 0, 4, 1, 0)  # sample-major is Whitted only
        with pytest.raises(RenderError):
            gs.set_launch_config(abi.KIND_WHITTED, 1, 0, 9, 1, 0)  # occupancy 1..8
