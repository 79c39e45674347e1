"""ft_render_whitted_device_noalloc (the allocation-free variant of
ft_render_whitted_device) against the entry point it mirrors, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from forgedthoughts_amd import synthetic
from forgedthoughts_amd.options import RaySettings, RenderOptions
from forgedthoughts_amd.render import GpuScene, library

from helpers import compile_src

pytestmark = pytest.mark.gpu

FT_ERR_INVALID_ARGUMENT = 5


def _noalloc(gs, opts, rs, rgb, hdr):
    o, s = opts.to_c(), rs.to_c()
    return library().ft_render_whitted_device_noalloc(gs.handle, C.addressof(o), C.addressof(s), rgb, hdr)


def test_noalloc_matches_whitted_device():
    """Two calls in a row, the second smaller, each bit-equal (bytes and radiance) to
    ft_render_whitted_device on the same scene; a request above the limit is refused."""
    import torch
    w = synthetic.whitted_1080p(width=160, height=96, spp=4)
    scene = compile_src(w.source)
    dev = torch.device("cuda:0")
    with GpuScene(scene, w.accel) as gs:
        for width, height, spp in ((160, 96, 4), (72, 40, 2)):
            print(f"[noalloc] {width}x{height} spp={spp}", flush=True)
            opts = RenderOptions(**{**w.options.__dict__, "width": width, "height": height})
            rs = RaySettings(4, 64, spp)
            n = width * height * 3
            ref_rgb = torch.zeros(n, dtype=torch.uint8, device=dev)
            ref_hdr = torch.zeros(n, dtype=torch.float32, device=dev)
            gs.render_whitted_device(opts, rs, ref_rgb.data_ptr(), ref_hdr.data_ptr())
            gs.synchronize()
            rgb = torch.full((n,), 7, dtype=torch.uint8, device=dev)
            hdr = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
            assert _noalloc(gs, opts, rs, rgb.data_ptr(), hdr.data_ptr()) == 0, library().ft_last_error()
            gs.synchronize()
            assert torch.equal(rgb, ref_rgb)
            assert torch.equal(hdr, ref_hdr)
            assert float(rgb.float().mean()) > 1.0
        big = RenderOptions(**{**w.options.__dict__, "width": 1920, "height": 1080})
        rc = _noalloc(gs, big, RaySettings(4, 64, 16), rgb.data_ptr(), None)
        assert rc == FT_ERR_INVALID_ARGUMENT
        assert b"allocation-free limit" in library().ft_last_error()


def test_noalloc_needs_a_built_kernel():
    """A scene whose specialised kernel was never built is refused (building it would load
    a code object), and renders once the ordinary entry point has built it."""
    import torch
    w = synthetic.whitted_1080p(width=64, height=32, spp=1)
    scene = compile_src(w.source)
    dev = torch.device("cuda:0")
    rgb = torch.zeros(64 * 32 * 3, dtype=torch.uint8, device=dev)
    with GpuScene(scene, w.accel) as gs:
        if not (gs.stats().flags & 1):
            pytest.skip("scene runs on the precompiled kernels")
        assert _noalloc(gs, w.options, w.settings, rgb.data_ptr(), None) == FT_ERR_INVALID_ARGUMENT
        gs.render_whitted_device(w.options, w.settings, rgb.data_ptr())
        gs.synchronize()
        ref = rgb.clone()
        rgb.zero_()
        assert _noalloc(gs, w.options, w.settings, rgb.data_ptr(), None) == 0
        gs.synchronize()
        assert torch.equal(rgb, ref)
