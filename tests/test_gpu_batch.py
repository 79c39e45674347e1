"""GPU batches (ft_batch_*) through the C ABI: every (scene, view) image of one batched
launch against the oracle rendering the same scene through the same camera, the batched
launch against the per-view launches of the single-scene path (identical bytes), and HIP
graph replays against the eager result."""
import numpy as np
import pytest

from forgedthoughts_amd import abi, synthetic
from forgedthoughts_amd.options import AccelMode, RaySettings, RenderOptions
from forgedthoughts_amd.render import GpuBatch
from forgedthoughts_amd.scene import RenderError
from oracle import ftref

from helpers import compile_src, with_camera

pytestmark = pytest.mark.gpu

BOTH = abi.BATCH_DEPTH | abi.BATCH_WHITTED


@pytest.fixture(scope="module")
def small_batch():
    w = synthetic.batch_256(n_scenes=6, views_per_scene=2, size=48, lo=8, hi=64, seed=9)
    scenes = [compile_src(s) for s in w.sources]
    return w, scenes


def test_batch_matches_oracle(small_batch):
    w, scenes = small_batch
    print(f"[batch parity] {len(scenes)} scenes x {len(w.views)} views {w.options.width}^2", flush=True)
    with GpuBatch(scenes, w.accel) as gb:
        views = gb.views(w.views)
        depth, whit = gb.render(views, w.options, w.settings, BOTH)
    for k, (sc, o, t, fov) in enumerate(w.views):
        ref = with_camera(scenes[sc], o, t, fov)
        rgb_c, t_c = ftref.render_depth(ref, w.accel, w.options, 1)
        d = np.abs(rgb_c.astype(int) - depth[k].astype(int)).max(axis=2)
        # hit/miss flips at grazing pixels move a byte from 0 to the shaded value
        assert (d > 1).mean() <= 1e-3 + 1.0 / d.size, f"view {k}: depth bytes differ on {(d > 1).sum()} px"
        rgb_c, _ = ftref.render_whitted(ref, w.accel, w.options, w.settings)
        d = np.abs(rgb_c.astype(int) - whit[k].astype(int)).max(axis=2)
        assert float((d <= 1).mean()) >= 0.999 - 1.0 / d.size, f"view {k}: whitted max {d.max()}"


def test_batched_equals_separate_launches(small_batch):
    w, scenes = small_batch
    with GpuBatch(scenes, w.accel) as gb:
        views = gb.views(w.views)
        d1, w1 = gb.render(views, w.options, w.settings, BOTH)
        d2, w2 = gb.render(views, w.options, w.settings, BOTH | abi.BATCH_SEPARATE)
        assert gb.stats().kernel_launches == 1 + 3 * len(w.views)
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(w1, w2)


def test_graph_replays_match_eager(small_batch):
    import torch
    w, scenes = small_batch
    n = len(w.views)
    bytes_ = n * w.options.width * w.options.height * 3
    with GpuBatch(scenes, w.accel) as gb:
        views = gb.views(w.views)
        eager_d, eager_w = gb.render(views, w.options, w.settings, BOTH)
        dd = torch.zeros(bytes_, dtype=torch.uint8, device="cuda")
        dw = torch.zeros(bytes_, dtype=torch.uint8, device="cuda")
        for rep in range(3):
            dd.zero_()
            dw.zero_()
            torch.cuda.synchronize()
            gb.render_device(views, w.options, w.settings, BOTH | abi.BATCH_GRAPH, dd.data_ptr(), dw.data_ptr())
            gb.synchronize()
            got_d = dd.cpu().numpy().reshape(eager_d.shape)
            got_w = dw.cpu().numpy().reshape(eager_w.shape)
            np.testing.assert_array_equal(got_d, eager_d, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(got_w, eager_w, err_msg=f"replay {rep}")
            # eager work between replays
            gb.render(views, w.options, w.settings, abi.BATCH_DEPTH)


def test_batch_errors(small_batch):
    w, scenes = small_batch
    with GpuBatch(scenes[:2], w.accel) as gb:
        with pytest.raises(RenderError):
            gb.render(gb.views([(5, (0, 0, 10), (0, 0, 0), 45.0)]), w.options, w.settings, BOTH)
        with pytest.raises(RenderError):
            gb.render(gb.views([(0, (0, 0, 10), (0, 0, 0), 45.0)]), w.options, w.settings, 0)


@pytest.mark.parametrize("shape", [(50, 50), (61, 37)], ids=["50x50", "61x37"])
def test_persistent_depth_and_primary_hit_reuse(shape):
    """The persistent depth pass (per-lane refill from chunks of tiles) and the Whitted pass
    that resumes from its primary hits (over each view's hits in tile order) give the bytes
    of the tiled launches, on image sizes that leave partial tiles in both directions and a
    partial last chunk."""
    import dataclasses
    w = synthetic.batch_256(n_scenes=5, views_per_scene=3, size=50, lo=8, hi=128, seed=21)
    opts = dataclasses.replace(w.options, width=shape[0], height=shape[1])
    scenes = [compile_src(s) for s in w.sources]
    print(f"[batch persistent] {len(scenes)} scenes x {len(w.views)} views {shape}", flush=True)
    with GpuBatch(scenes, w.accel) as gb:
        views = gb.views(w.views)
        d1, w1 = gb.render(views, opts, w.settings, BOTH)
        d2, w2 = gb.render(views, opts, w.settings, BOTH | abi.BATCH_TILED)
        _, w3 = gb.render(views, opts, w.settings, abi.BATCH_WHITTED)
        d4, _ = gb.render(views, opts, w.settings, abi.BATCH_DEPTH)
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(d1, d4)
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(w1, w3)
