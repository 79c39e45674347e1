"""Differentiable sphere-tree rendering over libft_hip.so (ft_diff_* entry points).

`DiffRenderer` wraps one parametric scene (n spheres in a balanced union_round tree) and
its views; `SphereTreeRender` is a torch.autograd.Function whose forward renders the
Newton-refined hit distance t* and the Lambert shade s for every view, and whose backward
runs the adjoint kernel (implicit differentiation through the march, FP64 reduction).
Parameter layout: [cx, cy, cz, r] per sphere, then the blend radius k of each join in
post order (5 * n - 1 values).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import abi
from .options import RenderOptions
from .render import _check, library

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = library()
    vp, u32 = C.c_void_p, C.c_uint32
    L.ft_diff_create.argtypes = [u32, vp, u32, vp, C.POINTER(C.c_float), C.c_float, C.POINTER(vp)]
    L.ft_diff_destroy.argtypes = [vp]
    L.ft_diff_param_count.argtypes = [vp, C.POINTER(u32)]
    L.ft_diff_set_params.argtypes = [vp, vp]
    L.ft_diff_get_params.argtypes = [vp, vp]
    L.ft_diff_forward.argtypes = [vp, vp, vp]
    L.ft_diff_forward_device.argtypes = [vp, vp, vp]
    L.ft_diff_synchronize.argtypes = [vp]
    L.ft_diff_backward_device.argtypes = [vp, vp, vp, vp]
    L.ft_diff_set_target.argtypes = [vp, vp, vp]
    L.ft_diff_loss_grad.argtypes = [vp, C.c_float, C.c_float, vp, vp]
    L.ft_diff_fit.argtypes = [vp, u32, C.c_float, C.c_float, C.c_float, vp]
    L.ft_diff_stats.argtypes = [vp, vp]
    L.ft_diff_set_flop_counting.argtypes = [vp, C.c_int]
    _bound = True


def param_count(n_prims: int) -> int:
    return 5 * n_prims - 1


class DiffRenderer:
    def __init__(self, n_prims: int, views: Sequence[tuple], options: RenderOptions,
                 light_dir=(0.4, 1.0, 0.3), normal_step: float = 1.0e-2):
        _bind()
        from .render import GpuBatch
        self.n_prims = n_prims
        self.n_params = param_count(n_prims)
        self.options = options
        self.n_views = len(views)
        self._views = GpuBatch.views(views)
        o = options.to_c()
        light = (C.c_float * 3)(*[float(x) for x in light_dir])
        self._h = C.c_void_p()
        _check(library().ft_diff_create(n_prims, C.addressof(self._views), len(views), C.addressof(o), light,
                                        float(normal_step), C.byref(self._h)))

    @property
    def shape(self):
        return (self.n_views, self.options.height, self.options.width)

    def close(self):
        if self._h:
            library().ft_diff_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.n_params
        _check(library().ft_diff_set_params(self._h, p.ctypes.data))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(library().ft_diff_get_params(self._h, p.ctypes.data))
        return p

    def forward(self):
        t = np.zeros(self.shape, np.float32)
        s = np.zeros(self.shape, np.float32)
        _check(library().ft_diff_forward(self._h, t.ctypes.data, s.ctypes.data))
        return t, s

    def forward_device(self, t_ptr: int = 0, s_ptr: int = 0) -> None:
        _check(library().ft_diff_forward_device(self._h, t_ptr or None, s_ptr or None))

    def synchronize(self) -> None:
        _check(library().ft_diff_synchronize(self._h))

    def backward_device(self, gt_ptr: int = 0, gs_ptr: int = 0) -> np.ndarray:
        g = np.zeros(self.n_params, np.float64)
        _check(library().ft_diff_backward_device(self._h, gt_ptr or None, gs_ptr or None, g.ctypes.data))
        return g

    def set_target(self, t_target, s_target) -> None:
        t = np.ascontiguousarray(t_target, np.float32)
        s = np.ascontiguousarray(s_target, np.float32)
        _check(library().ft_diff_set_target(self._h, t.ctypes.data, s.ctypes.data))

    def loss_grad(self, w_t: float = 1.0, w_s: float = 1.0):
        loss = C.c_double(0.0)
        g = np.zeros(self.n_params, np.float64)
        _check(library().ft_diff_loss_grad(self._h, w_t, w_s, C.byref(loss), g.ctypes.data))
        return loss.value, g

    def fit(self, steps: int, lr: float, w_t: float = 1.0, w_s: float = 1.0) -> np.ndarray:
        losses = np.zeros(steps, np.float64)
        _check(library().ft_diff_fit(self._h, steps, lr, w_t, w_s, losses.ctypes.data))
        return losses

    def stats(self) -> abi.RenderStatsC:
        st = abi.RenderStatsC()
        _check(library().ft_diff_stats(self._h, C.addressof(st)))
        return st

    def set_flop_counting(self, enable: bool) -> None:
        """Later forwards / adjoints run the *_counted kernels; stats().sdf_flops sums them."""
        _check(library().ft_diff_set_flop_counting(self._h, 1 if enable else 0))


def _autograd():
    import torch

    class SphereTreeRender(torch.autograd.Function):
        """params (float32, 5n-1) -> (t*, s) images [views, H, W] on the GPU."""

        @staticmethod
        def forward(ctx, params, renderer: DiffRenderer):
            renderer.set_params(params.detach().to("cpu", torch.float32).numpy())
            t = torch.empty(renderer.shape, dtype=torch.float32, device="cuda")
            s = torch.empty(renderer.shape, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            renderer.forward_device(t.data_ptr(), s.data_ptr())
            renderer.synchronize()
            ctx.renderer = renderer
            ctx.param_device = params.device
            return t, s

        @staticmethod
        def backward(ctx, gt, gs):
            gt = gt.contiguous().float() if gt is not None else None
            gs = gs.contiguous().float() if gs is not None else None
            torch.cuda.synchronize()
            g = ctx.renderer.backward_device(gt.data_ptr() if gt is not None else 0,
                                             gs.data_ptr() if gs is not None else 0)
            return torch.from_numpy(g).to(ctx.param_device, torch.float32), None

    return SphereTreeRender


_fn = None


def sphere_tree_render(params, renderer: DiffRenderer):
    """torch entry point: (t*, s) = sphere_tree_render(params, renderer)."""
    global _fn
    if _fn is None:
        _fn = _autograd()
    return _fn.apply(params, renderer)
