"""Pins the CPU oracle (oracle/ftref.cpp) before it is trusted as the checker.

What pins what:
  * the reference's own image fixtures (tests/fixtures/render_regression, copied to
    tests/golden/render_regression) with the reference test's tolerance
    (render_regression.rs:138-145) — march, shading, tone mapping end to end;
  * the reference's known-answer tests: SoftBlob distance/bounds (renderer.rs:9418-9465,
    lib.rs:1768-1815) and the distance_post hook value (lib.rs:1868-1920);
  * closed forms: primitive SDFs, hg_sdf identities, the ray/sphere hit distance, the
    rotation convention of to_local (renderer.rs:6216-6240), Lambert under a point light
    (lambert.rs:7-13, renderer.rs:4217-4276), the Fresnel and IOR debug AOVs
    (ray.rs:224-285), the depth and Reinhard byte mappings (renderer.rs:5215-5239);
  * invariants: identical hit masks over naive/bvh/bricks, thread-count determinism, and
    the f32 march against the same march in f64;
  * the differentiable sphere-tree forward (no reference counterpart): the analytic
    ray/sphere root for one sphere, f32 against f64, and smoothness of the f64 forward in
    the parameters (central differences of two step sizes agree), which is what makes its
    finite differences the gradient reference.
"""
import math
import os

import numpy as np
import pytest

from forgedthoughts_amd import synthetic
from forgedthoughts_amd.options import AccelMode, RaySettings, RenderOptions
from oracle import ftref

from helpers import REGRESSION, REGRESSION_CASES, compile_ft, compile_src, diff_metrics, read_png

F32 = np.float32


# ----------------------------------------------------------------- reference fixtures
@pytest.mark.parametrize("case", REGRESSION_CASES, ids=[c[0] for c in REGRESSION_CASES])
def test_reference_fixture(case):
    name, kind, opts, settings = case
    scene = compile_ft(os.path.join(REGRESSION, name + ".ft"))
    expected = read_png(os.path.join(REGRESSION, name + ".png"))
    if kind == "preview":
        rgb, _ = ftref.render_depth(scene, AccelMode.Bvh, opts, 1)
        # the depth path reproduces the reference bytes exactly
        assert diff_metrics(rgb, expected) == (0, 0.0)
    else:
        rgb, _ = ftref.render_whitted(scene, AccelMode.Bvh, opts, settings)
        mx, mean = diff_metrics(rgb, expected)
        assert mx <= 8 and mean <= 0.35, f"{name}: max {mx} mean {mean:.4f}"


# ------------------------------------------------------------------ known answers
SOFT_BLOB = """
sdf SoftBlob {
  let wave_scale = 0.15;
  fn bounds() { return vec3(1.2, 1.2, 1.0); }
  fn warp(p) { return vec3(p.x, p.y + sin(p.x * 4.0) * wave_scale, p.z); }
  fn distance(p) { let q = warp(p); return length(q) - 1.0; }
};
let scene = SoftBlob { material: Lambert { color: vec3(0.7, 0.45, 0.3) } };
"""


def test_soft_blob_known_answers():
    """renderer.rs:9458-9464: centre < -0.9, |surface| < 1e-4, bounds.max.x < 3;
    lib.rs:1809: distance(0, 1, 0) == 0."""
    scene = compile_src(SOFT_BLOB)
    pts = np.array([[0, 0, 0], [0, 1, 0]], np.float32)
    d, _, _ = ftref.distance_info(scene, AccelMode.Naive, pts)
    assert d[0] < -0.9
    assert abs(d[1]) < 1.0e-4
    assert d[1] == 0.0
    mn, mx, _ = ftref.scene_bounds(scene)
    assert mx[0] < 3.0
    fn = int(scene.nodes["fn_distance"][scene.root])
    out = ftref.eval_function(scene, fn, np.array([[0, 1, 0]], np.float32))
    assert out[0, 0] == 0.0


TWISTED_SHELL = """
sdf TwistedShell {
  let radius = 0.5;
  let thickness = 0.05;
  fn domain(p) { return rotate_y(p, p.y * 90.0); }
  fn distance(p) { return length(p) - radius; }
  fn distance_post(d, p) { return abs(d) - thickness; }
};
let scene = TwistedShell {};
"""


def test_distance_post_known_answer():
    """lib.rs:1901-1919: distance_post(0.2, (0,0,0)) == 0.15 (in f32)."""
    scene = compile_src(TWISTED_SHELL)
    fn = int(scene.nodes["fn_post"][scene.root])
    out = ftref.eval_function(scene, fn, np.array([[0.2, 0.2, 0.2, 0, 0, 0]], np.float32))
    assert out[0, 0] == F32(0.15)
    # domain(1, 1, 0) rotates by 90 degrees about y: (x, z) = (cos, -sin) * 1 -> (0, -1)
    fd = int(scene.nodes["fn_domain"][scene.root])
    q = ftref.eval_function(scene, fd, np.array([[1, 1, 0]], np.float32))[0]
    assert abs(q[0]) < 1e-6 and q[1] == 1.0 and abs(q[2] + 1.0) < 1e-6


# ------------------------------------------------------------------ closed forms
def _points(n, lo, hi, seed):
    return np.random.default_rng(seed).uniform(lo, hi, size=(n, 3)).astype(np.float32)


def test_sphere_and_box_sdf_closed_forms():
    scene = compile_src("var s = Sphere { radius: 0.75 }; s.pos.x = 1.0; s.pos.y = -0.5;\n"
                        "let scene = s;")
    p = _points(2000, -3, 3, 1)
    d, _, _ = ftref.distance_info(scene, AccelMode.Naive, p)
    q = p.astype(np.float64) - np.array([1.0, -0.5, 0.0])
    exact = np.linalg.norm(q, axis=1) - 0.75
    assert np.abs(d - exact).max() < 2e-6
    # Box.size is the full size (renderer.rs:1434-1435); IQ box distance
    scene = compile_src("let scene = Box { size: vec3(1.0, 2.0, 3.0) };")
    d, _, _ = ftref.distance_info(scene, AccelMode.Naive, p)
    h = np.array([0.5, 1.0, 1.5])
    a = np.abs(p.astype(np.float64)) - h
    exact = np.linalg.norm(np.maximum(a, 0), axis=1) + np.minimum(a.max(axis=1), 0)
    assert np.abs(d - exact).max() < 2e-6


def test_hg_sdf_identities():
    """Away from the blend band (both distances >= r) the smooth unions are plain min
    (renderer.rs:6019-6137); Subtract is max(a, -b)."""
    base = ("var a = Sphere { radius: 1.0 }; a.pos.x = -2.0;\n"
            "var b = Box { size: vec3(1.0, 1.0, 1.0) }; b.pos.x = 2.0;\n")
    p = _points(4000, -5, 5, 2)
    da, _, _ = ftref.distance_info(compile_src(base + "let scene = a;"), AccelMode.Naive, p)
    db, _, _ = ftref.distance_info(compile_src(base + "let scene = b;"), AccelMode.Naive, p)
    far = (da >= 0.3) & (db >= 0.3)
    for op in ("union_round", "union_chamfer", "union_soft"):
        d, _, _ = ftref.distance_info(compile_src(base + f"let scene = a.{op}(b, 0.3);"), AccelMode.Naive, p)
        np.testing.assert_array_equal(d[far & (np.abs(da - db) >= 0.3 * math.sqrt(2))],
                                      np.minimum(da, db)[far & (np.abs(da - db) >= 0.3 * math.sqrt(2))])
    d, _, _ = ftref.distance_info(compile_src(base + "let scene = a + b;"), AccelMode.Naive, p)
    np.testing.assert_array_equal(d, np.minimum(da, db))
    d, _, _ = ftref.distance_info(compile_src(base + "let scene = a - b;"), AccelMode.Naive, p)
    np.testing.assert_array_equal(d, np.maximum(da, -db))
    d, _, _ = ftref.distance_info(compile_src(base + "let scene = a & b;"), AccelMode.Naive, p)
    np.testing.assert_array_equal(d, np.maximum(da, db))


def test_rotation_convention():
    """to_local applies Rz(-z), Ry(-y), Rx(-x) (renderer.rs:6216-6240), so a shape with
    rot.y = 30 has the local +x edge centre of a triangle prism at (cos30, 0, -sin30)."""
    scene = compile_src("var t = ExtrudePolygon { sides: 3, radius: 1.0, height: 10.0 }; t.rot.y = 30.0;\n"
                        "let scene = t;")
    s = 2.0
    ang = math.radians(30.0)
    p = np.array([[s * math.cos(ang), 0.0, -s * math.sin(ang)],
                  [s * math.cos(ang), 0.0, s * math.sin(ang)]], np.float32)
    d, _, _ = ftref.distance_info(scene, AccelMode.Naive, p)
    apothem = 1.0 * math.cos(math.pi / 3)
    assert abs(d[0] - (s - apothem)) < 1e-5             # edge centre
    assert abs(d[1] - (s * math.cos(math.radians(60)) - apothem)) < 1e-5  # toward a vertex
    # rot.x = 90 swaps the y and z extents of a box
    a = compile_src("var b = Box { size: vec3(1.0, 2.0, 3.0) }; b.rot.x = 90.0; let scene = b;")
    b = compile_src("let scene = Box { size: vec3(1.0, 3.0, 2.0) };")
    pts = _points(500, -3, 3, 3)
    da, _, _ = ftref.distance_info(a, AccelMode.Naive, pts)
    db, _, _ = ftref.distance_info(b, AccelMode.Naive, pts)
    assert np.abs(da - db).max() < 1e-5


def _axis_camera(extra, width=9, height=9, fov=20.0):
    return ("let camera = Camera { origin: vec3(0.0, 0.0, 5.0), target: vec3(0.0, 0.0, 0.0), "
            f"fov_y: {fov} }};\n" + extra)


def test_ray_sphere_hit_distance():
    """Centre pixel of an odd-sized image looks down the axis: t = 5 - r - z_c."""
    for r, zc in ((1.0, 0.0), (0.5, 1.25), (2.0, -1.0)):
        src = _axis_camera(f"var s = Sphere {{ radius: {r} }}; s.pos.z = {zc}; let scene = s;")
        scene = compile_src(src)
        opts = RenderOptions(9, 9, 256, 40.0, 1.0e-4, 0.7, 6.0, 20.0)
        rgb, t = ftref.render_depth(scene, AccelMode.Naive, opts, 1)
        exact = 5.0 - r - zc
        assert abs(t[4, 4] - exact) < 1e-4
        # depth byte (renderer.rs:5215-5227): trunc(255 * min(((1 - clamp(t/far))^2)^(1/2.2), 1))
        v = (1.0 - min(max(t[4, 4] / 40.0, 0.0), 1.0)) ** 2
        assert rgb[4, 4, 0] == int(255.0 * min(v ** (1 / 2.2), 1.0))


def test_lambert_point_light_closed_form():
    """A Lambert wall facing the camera lit by a point light on the axis: radiance =
    color/pi * intensity / d^2 (lambert.rs:7-13, 1/d^2 falloff render_api.rs:171-289),
    unshadowed; then Reinhard + gamma for the byte (renderer.rs:5229-5239)."""
    src = _axis_camera("let key = PointLight { position: vec3(0.0, 0.0, 3.0), color: #ffffff, intensity: 9.0 };\n"
                       "let wall_mat = Lambert { color: #808080 };\n"
                       "var wall = Box { size: vec3(20.0, 20.0, 1.0), material: wall_mat }; wall.pos.z = -0.5;\n"
                       "let scene = wall;")
    scene = compile_src(src)
    opts = RenderOptions(9, 9, 256, 40.0, 1.0e-4, 0.7, 6.0, 20.0)
    rgb, hdr = ftref.render_whitted(scene, AccelMode.Bvh, opts, RaySettings(1, 9, 1))
    expect = (128.0 / 255.0) / math.pi * 9.0 / 9.0
    assert abs(hdr[4, 4, 0] - expect) / expect < 2e-4
    x = float(hdr[4, 4, 0])
    assert rgb[4, 4, 0] == int(255.0 * min((x / (1.0 + x)) ** (1 / 2.2), 1.0))


def test_fresnel_and_ior_aovs():
    """Debug AOVs at normal incidence on a dielectric (ray.rs:252-277):
    fresnel = ((n-1)/(n+1))^2, ior AOV = (n-1)/2."""
    src = _axis_camera("let glass = Dielectric { ior: 1.5 };\n"
                       "var wall = Box { size: vec3(20.0, 20.0, 1.0), material: glass }; wall.pos.z = -0.5;\n"
                       "let scene = wall;")
    scene = compile_src(src)
    opts = RenderOptions(9, 9, 256, 40.0, 1.0e-4, 0.7, 6.0, 20.0)
    _, hdr = ftref.render_whitted(scene, AccelMode.Bvh, opts, RaySettings(2, 9, 1, "fresnel"))
    assert abs(hdr[4, 4, 0] - 0.04) < 1e-5
    _, hdr = ftref.render_whitted(scene, AccelMode.Bvh, opts, RaySettings(2, 9, 1, "ior"))
    assert abs(hdr[4, 4, 0] - 0.25) < 1e-6


# ------------------------------------------------------------------- invariants
def test_accel_modes_agree_on_hits():
    w = synthetic.depth_512()
    scene = compile_src(w.source)
    opts = RenderOptions(**{**w.options.__dict__, "width": 96, "height": 96})
    ts = [ftref.render_depth(scene, a, opts, 1)[1] for a in (AccelMode.Naive, AccelMode.Bvh, AccelMode.Bricks)]
    for t in ts[1:]:
        np.testing.assert_array_equal(t >= 0, ts[0] >= 0)
        both = (t >= 0) & (ts[0] >= 0)
        assert np.abs(t[both] - ts[0][both]).max() <= 1e-4


def test_thread_count_determinism():
    w = synthetic.whitted_1080p(width=64, height=36, spp=1)
    scene = compile_src(w.source)
    a, ha = ftref.render_whitted(scene, w.accel, w.options, w.settings, threads=1)
    b, hb = ftref.render_whitted(scene, w.accel, w.options, w.settings, threads=8)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ha, hb)


def test_f32_march_against_f64():
    """The error of the f32 march (the reference's arithmetic) against the same algorithm
    in f64 on the 512x512 depth scene at 128x128: hit masks agree except grazing pixels,
    and hit distances agree to the march tolerance."""
    w = synthetic.depth_512()
    scene = compile_src(w.source)
    opts = RenderOptions(**{**w.options.__dict__, "width": 128, "height": 128})
    _, t32 = ftref.render_depth(scene, w.accel, opts, 1, precision=0)
    _, t64 = ftref.render_depth(scene, w.accel, opts, 1, precision=1)
    assert int(((t32 >= 0) != (t64 >= 0)).sum()) <= 2
    both = (t32 >= 0) & (t64 >= 0)
    assert np.abs(t32[both] - t64[both]).max() < 2 * w.options.epsilon


# ------------------------------------------------- differentiable sphere-tree forward
def _sphere_ray_t(o, d, c, r):
    oc = np.asarray(o, np.float64) - np.asarray(c, np.float64)
    b = float(np.dot(oc, d))
    disc = b * b - (float(np.dot(oc, oc)) - r * r)
    return -b - math.sqrt(disc) if disc >= 0 else None


def test_diff_forward_single_sphere_closed_form():
    """One sphere: t* is the analytic ray/sphere root; the centre pixel looks along -z."""
    views = [(0, (0.0, 0.0, 6.0), (0.0, 0.0, 0.0), 30.0)]
    opts = RenderOptions(9, 9, 256, 30.0, 1.0e-5, 0.7, 6.0, 30.0)
    for c, r in (((0.0, 0.0, 0.0), 1.0), ((0.0, 0.0, 0.5), 0.7), ((0.0, 0.0, -1.0), 1.5)):
        t, s = ftref.diff_forward(1, list(c) + [r], views, opts, (0.0, 0.0, 1.0), 1.0e-2, precision=1)
        exact = _sphere_ray_t((0, 0, 6), (0, 0, -1), c, r)
        assert abs(t[0, 4, 4] - exact) < 1e-9
        assert abs(s[0, 4, 4] - 1.0) < 1e-6   # normal (0,0,1) facing the light


def test_diff_forward_f32_against_f64_and_smoothness():
    w = synthetic.adjoint_fit(n_prims=8, n_views=2, size=48)
    t64, s64 = ftref.diff_forward(w.n_prims, w.params_true, w.views, w.options, w.light_dir, w.normal_step, precision=1)
    t32, s32 = ftref.diff_forward(w.n_prims, w.params_true, w.views, w.options, w.light_dir, w.normal_step, precision=0)
    assert int(((t64 >= 0) != (t32 >= 0)).sum()) <= 2
    m = (t64 >= 0) & (t32 >= 0)
    assert np.abs(t64 - t32)[m].max() < 1e-4 and np.abs(s64 - s32)[m].max() < 1e-3
    # the f64 forward is smooth in the parameters on interior pixels, so central
    # differences of step 1e-4 and 1e-5 agree: it can serve as the gradient reference
    from scipy.ndimage import binary_erosion
    interior = binary_erosion(t64 >= 0, structure=np.ones((1, 3, 3)))
    p = np.array(w.params_true, np.float64)

    def loss(q):
        t, s = ftref.diff_forward(w.n_prims, q, w.views, w.options, w.light_dir, w.normal_step, precision=1)
        return float(((t * interior) ** 2).sum() + ((s * interior) ** 2).sum())

    for j in (0, 3, 4 * w.n_prims):
        e = np.zeros_like(p)
        e[j] = 1.0
        g4 = (loss(p + 1e-4 * e) - loss(p - 1e-4 * e)) / 2e-4
        g5 = (loss(p + 1e-5 * e) - loss(p - 1e-5 * e)) / 2e-5
        assert abs(g4 - g5) <= 1e-4 * max(abs(g4), 1.0), (j, g4, g5)


# -------------------------------------------------------------- material hooks
# Per-hit hooks of a material definition (resolve_leaf_material renderer.rs:7164-7536, the
# arms for models other than Standard / Layered) evaluated with the shading context of
# make_shading_context (renderer.rs:7875-7911).  A black, unlit surface whose emission
# colour is a hook of the context radiates exactly that colour, so the HDR output of a
# one-bounce Whitted render reads the context fields back.
_DARK = "let key = PointLight { position: vec3(0.0, 0.0, 3.0), color: #ffffff, intensity: 0.0 };\n"
_OPTS9 = RenderOptions(9, 9, 256, 40.0, 1.0e-4, 0.7, 6.0, 20.0)


def _emitter(expr, shape="Sphere { radius: 1.0, material: E {} }", place=""):
    return compile_src(_axis_camera(
        _DARK + "material E { model: Lambert; color = vec3(0.0, 0.0, 0.0); emission_strength = 1.0;\n"
        f"  fn emission_color(ctx) {{ return {expr}; }} }};\n" f"var s = {shape}; {place}\nlet scene = s;"))


def test_hook_context_normal_matches_normal_aov():
    scene = _emitter("ctx.normal * 0.5 + vec3(0.5, 0.5, 0.5)")
    _, hdr = ftref.render_whitted(scene, AccelMode.Bvh, _OPTS9, RaySettings(1, 9, 1))
    _, aov = ftref.render_whitted(scene, AccelMode.Bvh, _OPTS9, RaySettings(1, 9, 1, "normal"))
    assert np.abs(hdr - aov).max() < 1e-6 and aov.std() > 0.01


def test_hook_context_points_and_view_direction():
    """Centre pixel of the axis camera hits a sphere of radius r at z_c on (0, 0, z_c + r)."""
    r, zc = 0.75, 0.5
    place = f"s.pos.z = {zc};"
    shape = f"Sphere {{ radius: {r}, material: E {{}} }}"
    for expr, want in (("ctx.position * 0.1 + vec3(0.5, 0.5, 0.5)", (0.5, 0.5, 0.5 + 0.1 * (zc + r))),
                       ("ctx.local_position * 0.1 + vec3(0.5, 0.5, 0.5)", (0.5, 0.5, 0.5 + 0.1 * r)),
                       ("ctx.view_dir * 0.25 + vec3(0.5, 0.5, 0.5)", (0.5, 0.5, 0.75))):
        scene = _emitter(expr, shape, place)
        _, hdr = ftref.render_whitted(scene, AccelMode.Bvh, _OPTS9, RaySettings(1, 9, 1))
        assert np.abs(hdr[4, 4] - np.array(want)).max() < 2e-4, expr


def test_hook_context_ids():
    scene = _emitter("vec3(ctx.front_face, ctx.object_id + 0.5, ctx.material_id + 0.25) * 0.5")
    leaf = int(np.flatnonzero(scene.nodes["kind"] == 0)[0])
    obj, mat = int(scene.nodes["object_id"][leaf]), int(scene.nodes["material_id"][leaf])
    _, hdr = ftref.render_whitted(scene, AccelMode.Bvh, _OPTS9, RaySettings(1, 9, 1))
    assert np.abs(hdr[4, 4] - 0.5 * np.array([1.0, obj + 0.5, mat + 0.25])).max() < 1e-6


def test_constant_standard_hooks_render_as_static_materials():
    """Constant hooks of a Standard-model definition give the image of the static definition
    with those values: metallic, specular, specular weight and transmission replace their
    parameters, the coat hooks replace the specular colour, ior and clearcoat roughness, and
    the clearcoat is the static coat weight times the coat mask (0.4 * 0.5 = 0.2); an ior hook
    is overridden by the coat ior (renderer.rs:7259-7470)."""
    lit = "let key = PointLight { position: vec3(0.5, 0.7, 3.0), color: #ffffff, intensity: 9.0 };\n"
    wall = "Box { size: vec3(20.0, 20.0, 1.0), material: M {} }"
    hooked = compile_src(_axis_camera(
        lit + "material M { model: Standard; color = #ffffff; coat_weight = 0.4;\n"
        "  fn color(ctx) { return vec3(0.25, 0.5, 0.75); }\n"
        "  fn metallic(ctx) { return 0.3; }  fn specular(ctx) { return 0.6; }\n"
        "  fn specular_weight(ctx) { return 0.8; }  fn transmission(ctx) { return 0.2; }\n"
        "  fn coat_color(ctx) { return vec3(0.9, 0.8, 0.7); }  fn coat_roughness(ctx) { return 0.2; }\n"
        "  fn coat_ior(ctx) { return 1.6; }  fn ior(ctx) { return 2.5; }  fn coat_mask(ctx) { return 0.5; } };\n"
        f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
    static = compile_src(_axis_camera(
        lit + "material M { model: Standard; color = vec3(0.25, 0.5, 0.75); metallic = 0.3;\n"
        "  specular = 0.6; specular_weight = 0.8; transmission = 0.2; coat_color = vec3(0.9, 0.8, 0.7);\n"
        "  coat_roughness = 0.2; coat_ior = 1.6; coat_weight = 0.2; };\n"
        f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
    assert len(hooked.material_hooks) == 1 and len(static.material_hooks) == 0
    h = hooked.material_hooks[0]
    assert h["ior"] >= 0 and h["clearcoat"] < 0 and h["coat_mask"] >= 0
    for aov in (None, "ior", "transmission", "fresnel"):
        a, ha = ftref.render_whitted(hooked, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1, aov))
        b, hb = ftref.render_whitted(static, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1, aov))
        assert np.array_equal(a, b) and np.abs(ha - hb).max() == 0.0 and ha.max() > 0.0, aov
    _, hdr = ftref.render_whitted(hooked, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1, "ior"))
    assert abs(hdr[4, 4, 0] - 0.3) < 1e-6
    # the image does depend on the clearcoat (an unmasked weight) and on the metallic term
    _, ha = ftref.render_whitted(hooked, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    for change in (("coat_weight = 0.2", "coat_weight = 0.4"), ("metallic = 0.3", "metallic = 0.0")):
        other = compile_src(_axis_camera(
            lit + "material M { model: Standard; color = vec3(0.25, 0.5, 0.75); metallic = 0.3;\n"
            "  specular = 0.6; specular_weight = 0.8; transmission = 0.2; coat_color = vec3(0.9, 0.8, 0.7);\n"
            "  coat_roughness = 0.2; coat_ior = 1.6; coat_weight = 0.2; };\n".replace(*change) +
            f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
        _, hc = ftref.render_whitted(other, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
        assert np.abs(ha - hc).max() > 1e-4, change
    # a coat weight hook without a mask is the clearcoat itself
    weight = compile_src(_axis_camera(
        lit + "material M { model: Standard; fn coat_weight(ctx) { return 0.2; } };\n"
        f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
    plain = compile_src(_axis_camera(
        lit + "material M { model: Standard; coat_weight = 0.2; };\n"
        f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
    a, ha = ftref.render_whitted(weight, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    b, hb = ftref.render_whitted(plain, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    assert np.abs(ha - hb).max() == 0.0


def test_bump_and_normal_hooks_reshape_the_shading_normal():
    """resolve_dynamic_bump / resolve_dynamic_normal (renderer.rs:7749-7827) on a wall facing
    +z: a constant bump and a normal hook returning the context normal or its negation leave
    the image of the static material; a bump linear in local x (0.5 x) tilts the normal by
    finite differences along the tangent frame ((-1,0,0), (0,-1,0)) to normalize(-0.5, 0, 1),
    the image of a normal hook returning that vector."""
    lit = "let key = PointLight { position: vec3(0.6, 0.4, 3.0), color: #ffffff, intensity: 9.0 };\n"
    wall = "var w = Box { size: vec3(20.0, 20.0, 1.0), material: M {} }; w.pos.z = -0.5; let scene = w;"

    def render(hooks):
        sc = compile_src(_axis_camera(lit + f"material M {{ model: Lambert; color = #c0c0c0; {hooks} }};\n" + wall))
        return ftref.render_whitted(sc, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))

    a, ha = render("")
    for hooks in ("fn bump(ctx) { return 0.3; }", "fn normal(ctx) { return ctx.normal; }",
                  "fn normal(ctx) { return ctx.normal * -2.0; }"):
        b, hb = render(hooks)
        assert np.array_equal(a, b) and np.abs(ha - hb).max() == 0.0, hooks
    b, hb = render("fn bump(ctx) { return ctx.local_position.x * 0.5; }")
    c, hc = render("fn normal(ctx) { return vec3(-0.5, 0.0, 1.0); }")
    assert np.abs(hb - hc).max() < 1e-4 * max(1.0, float(hc.max()))
    assert np.abs(ha - hb).max() > 1e-3


def test_medium_hook_matches_a_static_medium():
    """A medium hook returning a constant Medium object gives the image of the static
    definition with that medium (medium_from_value renderer.rs:6834-6845: ior clamped to
    [1, 3], absorption colour, density >= 0; missing fields take 1, white, 0), and the image
    depends on it (the refracted rays cross the absorbing medium, transition_medium)."""
    lit = "let key = PointLight { position: vec3(0.0, 2.0, 4.0), color: #ffffff, intensity: 20.0 };\n"
    scene = ("var w = Box { size: vec3(20.0, 20.0, 1.0), material: Lambert { color: #d0c0a0 } }; w.pos.z = -3.0;\n"
             "let scene = w + Sphere { radius: 1.0, material: M {} };")

    def render(body):
        sc = compile_src(_axis_camera(lit + f"material M {{ model: Dielectric; ior = 1.5; {body} }};\n" + scene))
        return sc, ftref.render_whitted(sc, AccelMode.Bvh, _OPTS9, RaySettings(4, 9, 1))

    hooked, (a, ha) = render("fn medium(ctx) { return Medium { ior: 1.4, absorption_color: vec3(0.5, 0.6, 0.7), "
                             "density: 0.3 }; }")
    assert len(hooked.material_hooks) == 1 and hooked.material_hooks[0]["medium_density"] >= 0
    _, (b, hb) = render("medium = Medium { ior: 1.4, absorption_color: vec3(0.5, 0.6, 0.7), density: 0.3 };")
    assert np.array_equal(a, b) and np.abs(ha - hb).max() == 0.0
    _, (c, hc) = render("")
    assert np.abs(ha - hc).max() > 1e-3
    _, (d, hd) = render("fn medium(ctx) { return Medium { density: 0.0 }; }")
    _, (e, he) = render("medium = Medium { ior: 1.0, absorption_color: vec3(1.0, 1.0, 1.0), density: 0.0 };")
    assert np.abs(hd - he).max() == 0.0


def test_eval_hook_replaces_the_bsdf_while_the_dynamic_id_lasts():
    """eval_bsdf (renderer.rs:4905-4920): a constant eval hook f gives f times the light sum
    (twice f, twice the image), close to the Lambert it stands in for; a lerped set has no
    dynamic id (blend_params renderer.rs:7624), so below the splitting smooth node the hook
    is not read even where the blend weight is 0."""
    lit = "let key = PointLight { position: vec3(0.5, 0.7, 3.0), color: #ffffff, intensity: 9.0 };\n"
    wall = "var w = Box { size: vec3(20.0, 20.0, 1.0), material: M {} }; w.pos.z = -0.5; let scene = w;"

    def render(body, scene=wall):
        sc = compile_src(_axis_camera(lit + f"material M {{ model: Lambert; color = vec3(0.25, 0.5, 0.75); {body} }};\n"
                                      + scene))
        return ftref.render_whitted(sc, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))[1]

    plain = render("")
    one = render("fn eval(ctx) { return vec3(0.25, 0.5, 0.75) * 0.31830988618; }")
    # the Lambert BSDF is colour / pi plus a small term of its own: the hook's image is close
    # to the plain one and not equal to it
    assert 0.0 < np.abs(one - plain).max() < 0.05 * float(plain.max())
    two = render("fn eval(ctx) { return vec3(0.25, 0.5, 0.75) * 0.63661977236; }")
    assert np.abs(two - 2.0 * one).max() <= 1e-5 * max(1.0, float(one.max()))   # f times the light sum
    # the first smooth node below the root splits the hit into its two child materials
    # (resolve_split_material_from_node), each keeping its id; a smooth node below that
    # blends, and its lerp has no id: the wall, far from both spheres, renders without the hook
    blended = ("var w = Box { size: vec3(20.0, 20.0, 1.0), material: M {} }; w.pos.z = -0.5;\n"
               "var s = Sphere { radius: 0.3, material: M {} }; s.pos.x = 6.0;\n"
               "var u = Sphere { radius: 0.3, material: M {} }; u.pos.x = -6.0;\n"
               "let scene = w.union_round(s, 0.05).union_round(u, 0.05);")
    assert np.abs(render("fn eval(ctx) { return vec3(5.0); }", blended) - render("", blended)).max() == 0.0
    split = ("var w = Box { size: vec3(20.0, 20.0, 1.0), material: M {} }; w.pos.z = -0.5;\n"
             "var s = Sphere { radius: 0.3, material: M {} }; s.pos.x = 6.0;\n"
             "let scene = w.union_round(s, 0.05);")
    assert np.abs(render("fn eval(ctx) { return vec3(5.0); }", split) - render("", split)).max() > 1.0


def test_hooks_outside_the_whitted_arms_leave_materials_static():
    """pdf / sample (path tracer only) and subsurface hooks, hooks a model's arm does not
    read (a Metal's ior) and helper functions are never called by the Whitted render: such a
    material compiles as its static parameters."""
    lit = "let key = PointLight { position: vec3(0.0, 0.0, 3.0), color: #ffffff, intensity: 9.0 };\n"
    wall = "var w = Box { size: vec3(20.0, 20.0, 1.0), material: M {} }; w.pos.z = -0.5; let scene = w;"
    hooked = compile_src(_axis_camera(
        lit + "material M { model: Metal; color = #c0a080; roughness = 0.3;\n"
        "  fn helper(ctx) { return ctx.position.x; }  fn ior(ctx) { return 2.0; }\n"
        "  fn pdf(ctx) { return helper(ctx); }  fn sample(ctx) { return 1.0; }\n"
        "  fn subsurface(ctx) { return 0.0; } };\n" + wall))
    static = compile_src(_axis_camera(lit + "material M { model: Metal; color = #c0a080; roughness = 0.3; };\n" + wall))
    assert len(hooked.material_hooks) == 0
    assert all(int(m["dynamic_material_id"]) < 0 for m in hooked.materials)
    a, ha = ftref.render_whitted(hooked, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    b, hb = ftref.render_whitted(static, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    assert np.array_equal(a, b) and np.abs(ha - hb).max() == 0.0 and ha.max() > 0.0


def test_constant_hooks_render_as_static_materials():
    """A hook returning a constant gives the image of the static material with that value
    (colour on a lit Lambert-model definition; ior read back by the IOR and Fresnel AOVs)."""
    lit = "let key = PointLight { position: vec3(0.0, 0.0, 3.0), color: #ffffff, intensity: 9.0 };\n"
    wall = "Box { size: vec3(20.0, 20.0, 1.0), material: M {} }"
    hooked = compile_src(_axis_camera(lit + "material M { model: Lambert; color = #ffffff;\n"
                                      "  fn color(ctx) { return vec3(0.25, 0.5, 0.75); } };\n"
                                      f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
    static = compile_src(_axis_camera(lit + "material M { model: Lambert; color = vec3(0.25, 0.5, 0.75); };\n"
                                      f"var w = {wall}; w.pos.z = -0.5; let scene = w;"))
    a, ha = ftref.render_whitted(hooked, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    b, hb = ftref.render_whitted(static, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1))
    assert np.array_equal(a, b) and np.abs(ha - hb).max() == 0.0 and ha.max() > 0.0
    glass = compile_src(_axis_camera("material G { model: Dielectric; fn ior(ctx) { return 1.5; } };\n"
                                     f"var w = {wall.replace('M {}', 'G {}')}; w.pos.z = -0.5; let scene = w;"))
    assert len(glass.material_hooks) == 1
    _, hdr = ftref.render_whitted(glass, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1, "ior"))
    assert abs(hdr[4, 4, 0] - 0.25) < 1e-6
    _, hdr = ftref.render_whitted(glass, AccelMode.Bvh, _OPTS9, RaySettings(2, 9, 1, "fresnel"))
    assert abs(hdr[4, 4, 0] - 0.04) < 1e-5
