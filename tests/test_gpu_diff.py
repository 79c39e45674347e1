"""GPU differentiable rendering (ft_diff_*) through the C ABI.

  * forward (t*, s) against the oracle's f32 forward of the same parametric scene;
  * adjoint gradient against central finite differences (h = 1e-4) of the oracle's FP64
    forward, every parameter, relative error <= 1e-3;
  * one sphere: dt*/d(centre, radius) against the analytic derivative of the ray/sphere root;
  * Adam fit lowers the loss; the torch.autograd.Function gives the C-ABI gradient.
"""
import math

import numpy as np
import pytest

from forgedthoughts_amd import synthetic
from forgedthoughts_amd.diff import DiffRenderer, sphere_tree_render
from forgedthoughts_amd.options import RenderOptions
from oracle import ftref

pytestmark = pytest.mark.gpu


def interior(t, max_jump=0.1):
    """Pixels whose 3x3 neighbourhood hits the surface with no depth discontinuity: away
    from silhouettes and occlusion edges, where t is a smooth function of the parameters."""
    from scipy.ndimage import binary_erosion, maximum_filter, minimum_filter
    hit = t >= 0
    out = binary_erosion(hit, structure=np.ones((1, 3, 3)), border_value=0)
    tt = np.where(hit, t, 0.0)
    jump = maximum_filter(tt, size=(1, 3, 3)) - minimum_filter(tt, size=(1, 3, 3))
    return out & (jump < max_jump)


@pytest.fixture(scope="module")
def small():
    return synthetic.adjoint_fit(n_prims=8, n_views=2, size=64, steps=30)


def test_forward_matches_oracle(small):
    w = small
    print(f"[diff forward] n={w.n_prims} views={len(w.views)} {w.options.width}^2", flush=True)
    with DiffRenderer(w.n_prims, w.views, w.options, w.light_dir, w.normal_step) as dr:
        dr.set_params(w.params_true)
        t, s = dr.forward()
    tc, sc = ftref.diff_forward(w.n_prims, np.float32(w.params_true), w.views, w.options, w.light_dir,
                                w.normal_step, precision=0)
    mism = int(((t >= 0) != (tc >= 0)).sum())
    assert mism <= max(1, int(1e-4 * t.size)), mism
    both = (t >= 0) & (tc >= 0)
    assert np.abs(t - tc)[both].max() <= 1e-4
    assert np.abs(s - sc)[both].max() <= 2e-3


def test_gradient_against_fp64_finite_differences(small):
    w = small
    p32 = np.float32(w.params_true) + np.float32(0.02) * np.float32(np.sin(np.arange(len(w.params_true))))
    p = p32.astype(np.float64)
    tt, st = ftref.diff_forward(w.n_prims, w.params_true, w.views, w.options, w.light_dir, w.normal_step, precision=1)
    import torch
    with DiffRenderer(w.n_prims, w.views, w.options, w.light_dir, w.normal_step) as dr:
        dr.set_params(p32)
        t, s = dr.forward()
        m = interior(t) & interior(tt)
        gt = torch.from_numpy((2.0 * (t - tt) * m).astype(np.float32)).cuda()
        gs = torch.from_numpy((2.0 * (s - st) * m).astype(np.float32)).cuda()
        torch.cuda.synchronize()
        g = dr.backward_device(gt.data_ptr(), gs.data_ptr())

    def loss(q):
        a, b = ftref.diff_forward(w.n_prims, q, w.views, w.options, w.light_dir, w.normal_step, precision=1)
        return float((((a - tt) ** 2) * m).sum() + (((b - st) ** 2) * m).sum())

    h = 1e-4
    fd = np.zeros_like(p)
    for j in range(p.size):
        e = np.zeros_like(p)
        e[j] = h
        fd[j] = (loss(p + e) - loss(p - e)) / (2 * h)
    err = np.abs(g - fd).max() / np.abs(fd).max()
    print(f"[diff grad] max |g - fd| / max |fd| = {err:.2e}", flush=True)
    assert err <= 1e-3, (g, fd)


def test_single_sphere_analytic_derivative():
    views = [(0, (0.3, 0.2, 6.0), (0.0, 0.0, 0.0), 30.0)]
    opts = RenderOptions(33, 33, 256, 30.0, 1.0e-5, 0.7, 6.0, 30.0)
    c = np.array([0.1, -0.2, 0.3])
    r = 1.1
    import torch
    with DiffRenderer(1, views, opts, (0.0, 0.0, 1.0), 1.0e-2) as dr:
        dr.set_params(np.float32(list(c) + [r]))
        t, _ = dr.forward()
        for (y, x) in ((16, 16), (10, 20), (22, 13)):
            seed = np.zeros_like(t)
            seed[0, y, x] = 1.0
            gt = torch.from_numpy(seed).cuda()
            torch.cuda.synchronize()
            g = dr.backward_device(gt.data_ptr(), 0)
            # the pixel's ray, from the camera basis of the view (origin o, target 0, fov 30)
            o = np.array(views[0][1])
            fwd = -o / np.linalg.norm(o)
            right = np.cross(fwd, [0, 1, 0]); right /= np.linalg.norm(right)
            up = np.cross(right, fwd); up /= np.linalg.norm(up)
            tanf = math.tan(math.radians(15.0))
            px = ((x + 0.5) / 33) * 2 - 1
            py = 1 - ((y + 0.5) / 33) * 2
            d = fwd + right * px * tanf + up * py * tanf
            d /= np.linalg.norm(d)
            hit = o + d * t[0, y, x]
            n = (hit - c) / np.linalg.norm(hit - c)
            # root of |o + t d - c| = r: dt/dr = 1 / (n . d), dt/dc = n / (n . d)
            nd = float(np.dot(n, d))
            dt_dr = 1.0 / nd
            dt_dc = n / nd
            assert abs(g[3] - dt_dr) <= 1e-3 * abs(dt_dr), (g, dt_dr)
            assert np.abs(g[:3] - dt_dc).max() <= 1e-3 * np.abs(dt_dc).max() + 1e-6, (g, dt_dc)


def test_fit_lowers_loss(small):
    w = small
    tt, st = ftref.diff_forward(w.n_prims, w.params_true, w.views, w.options, w.light_dir, w.normal_step, precision=0)
    with DiffRenderer(w.n_prims, w.views, w.options, w.light_dir, w.normal_step) as dr:
        dr.set_params(w.params_init)
        dr.set_target(tt.astype(np.float32), st.astype(np.float32))
        losses = dr.fit(w.steps, w.lr)
        assert np.isfinite(losses).all()
        assert losses[-1] < 0.5 * losses[0], losses
        p = dr.get_params()
    e0 = np.abs(np.float32(w.params_init) - np.float32(w.params_true)).mean()
    e1 = np.abs(p - np.float32(w.params_true)).mean()
    print(f"[diff fit] loss {losses[0]:.4g} -> {losses[-1]:.4g}; param error {e0:.4f} -> {e1:.4f}", flush=True)


def test_torch_autograd_function(small):
    import torch
    w = small
    tt, st = ftref.diff_forward(w.n_prims, w.params_true, w.views, w.options, w.light_dir, w.normal_step, precision=0)
    with DiffRenderer(w.n_prims, w.views, w.options, w.light_dir, w.normal_step) as dr:
        params = torch.tensor(w.params_init, dtype=torch.float32, device="cuda", requires_grad=True)
        t, s = sphere_tree_render(params, dr)
        m = torch.from_numpy(interior(t.detach().cpu().numpy()) & interior(tt)).cuda()
        T = torch.from_numpy(tt.astype(np.float32)).cuda()
        S = torch.from_numpy(st.astype(np.float32)).cuda()
        loss = (((t - T) ** 2) * m).sum() + (((s - S) ** 2) * m).sum()
        loss.backward()
        gt = (2 * (t.detach() - T) * m).contiguous()
        gs = (2 * (s.detach() - S) * m).contiguous()
        torch.cuda.synchronize()
        ref = dr.backward_device(gt.data_ptr(), gs.data_ptr())
    np.testing.assert_allclose(params.grad.cpu().numpy(), np.float32(ref), rtol=1e-5, atol=1e-6)
