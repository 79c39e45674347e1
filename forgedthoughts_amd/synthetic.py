"""Seeded synthetic scenes for the benchmark workloads, emitted as Forge source.

Each generator returns Forge text (parsed by the same front end as any `.ft` file) plus
the render settings of the workload.  Randomness: SplitMix64 seeded with the workload
index, so every scene is reproducible bit for bit.

Workloads
  depth_512        512x512 depth preview, naive accel: ground box (top face at y=-1), sphere
                   r=1 union_round a 0.5 box at x=1.5 (r=0.1); camera (0,1,-4)->origin, fov 45;
                   256 steps, eps 1e-4, far 100
  whitted_1080p    1920x1080 Whitted, 16 spp, max depth 4, BVH accel: 64 primitives
                   (sphere/box/torus, centres U[-5,5]^3, sizes U[0.2,0.8]) in a balanced
                   union_round (r=0.1) tree; Lambert / Metal (roughness U[0,0.4]) /
                   Dielectric (ior 1.52); PointLight + EnvLight + sky environment
  bricks_4096      4096 primitives in a balanced hard-union tree plus a twisted statue
                   custom SDF (domain twist rotate_y(p, p.y*18), ripple distance_post)
  batch_256        256 scenes with log-uniform primitive counts in [8, 2048] and mixed
                   union/diff/intersect x round/chamfer/stairs operators
  adjoint_32       32 primitives with position/radius/blend parameters
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

from .options import AccelMode, RaySettings, RenderOptions

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def choice(self, n: int) -> int:
        return self.next_u64() % n


def _f(x: float) -> str:
    s = f"{x:.6f}"
    return s if not s.startswith("-") else f"(0.0 - {s[1:]})"


def _vec(x, y, z) -> str:
    return f"vec3({_f(x)}, {_f(y)}, {_f(z)})"


@dataclass
class Workload:
    name: str
    source: str
    options: RenderOptions
    accel: AccelMode
    settings: Optional[RaySettings] = None
    aa: int = 1
    kind: str = "depth"   # "depth" or "whitted"


def depth_512() -> Workload:
    src = f"""
let camera = Camera {{ origin: {_vec(0, 1, -4)}, target: {_vec(0, 0, 0)}, fov_y: 45.0 }};
var ground = Box {{ size: {_vec(200, 1, 200)} }};
ground.pos.y = -1.5;
var ball = Sphere {{ radius: 1.0 }};
var cube = Box {{ size: {_vec(0.5, 0.5, 0.5)} }};
cube.pos.x = 1.5;
let scene = ground + ball.union_round(cube, 0.1);
"""
    opts = RenderOptions(width=512, height=512, max_steps=256, max_dist=100.0, epsilon=1.0e-4,
                         step_scale=0.7, camera_z=6.0, fov_y_degrees=45.0)
    return Workload("depth_512", src, opts, AccelMode.Naive, None, 1, "depth")


SKY = """
environment Sky {
  let zenith = #4a86e8;
  let horizon = #dcebff;
  let nadir = #eef2f8;
  fn color(dir) {
    let t = clamp(dir.y * 0.5 + 0.5, 0.0, 1.0);
    let sky = mix(horizon, zenith, pow(t, 0.8));
    let down = clamp(-dir.y * 1.6, 0.0, 1.0);
    return mix(sky, nadir, down);
  }
};
"""


def _random_primitive(rng: SplitMix64, name: str, mat_name: str) -> str:
    kind = rng.choice(3)
    c = [rng.uniform(-5.0, 5.0) for _ in range(3)]
    s = rng.uniform(0.2, 0.8)
    if kind == 0:
        body = f"Sphere {{ radius: {_f(s)}, material: {mat_name} }}"
    elif kind == 1:
        body = f"Box {{ size: {_vec(s, s, s)}, material: {mat_name} }}"
    else:
        body = f"Torus {{ major_radius: {_f(s)}, minor_radius: {_f(s / 4.0)}, material: {mat_name} }}"
    return (f"var {name} = {body};\n{name}.pos.x = {_f(c[0])};\n{name}.pos.y = {_f(c[1])};\n"
            f"{name}.pos.z = {_f(c[2])};\n")


def _random_material(rng: SplitMix64, name: str) -> str:
    kind = rng.choice(3)
    col = [rng.uniform(0.2, 0.9) for _ in range(3)]
    if kind == 0:
        return f"let {name} = Lambert {{ color: {_vec(*col)} }};\n"
    if kind == 1:
        return f"let {name} = Metal {{ color: {_vec(*col)}, roughness: {_f(rng.uniform(0.0, 0.4))} }};\n"
    return f"let {name} = Dielectric {{ ior: 1.52 }};\n"


def balanced(names: List[str], op: str, r: float) -> str:
    if len(names) == 1:
        return names[0]
    mid = len(names) // 2
    lhs, rhs = balanced(names[:mid], op, r), balanced(names[mid:], op, r)
    if op == "+":
        return f"({lhs} + {rhs})"
    return f"{lhs}.{op}({rhs}, {_f(r)})"


def whitted_1080p(n_prims: int = 64, width: int = 1920, height: int = 1080, spp: int = 16,
                  max_depth: int = 4, seed: int = 1) -> Workload:
    rng = SplitMix64(seed)
    parts = [f"let camera = Camera {{ origin: {_vec(0, 2, -14)}, target: {_vec(0, 0, 0)}, fov_y: 45.0 }};\n",
             f"let key = PointLight {{ position: {_vec(6, 9, -8)}, color: #fff2d8, intensity: 180.0 }};\n",
             "let skylight = EnvLight { color: #cfe0ff, intensity: 0.25 };\n", SKY]
    names = []
    for i in range(n_prims):
        parts.append(_random_material(rng, f"m{i}"))
        parts.append(_random_primitive(rng, f"p{i}", f"m{i}"))
        names.append(f"p{i}")
    parts.append(f"let scene = {balanced(names, 'union_round', 0.1)};\n")
    opts = RenderOptions(width=width, height=height, max_steps=256, max_dist=40.0, epsilon=1.0e-3,
                         step_scale=0.7, camera_z=6.0, fov_y_degrees=45.0)
    settings = RaySettings(max_depth=max_depth, tile_size=64, aa_samples=spp)
    return Workload("whitted_1080p", "".join(parts), opts, AccelMode.Bvh, settings, spp, "whitted")


def bricks_4096(n_prims: int = 4096, width: int = 1920, height: int = 1080, seed: int = 2) -> Workload:
    rng = SplitMix64(seed)
    parts = [f"let camera = Camera {{ origin: {_vec(0, 2, -16)}, target: {_vec(0, 0, 0)}, fov_y: 45.0 }};\n",
             f"let key = PointLight {{ position: {_vec(6, 9, -8)}, color: #fff2d8, intensity: 180.0 }};\n",
             SKY,
             """
sdf TwistedStatue {
  let width = 0.55;
  let height = 1.5;
  let depth = 0.42;
  let twist = 18.0;
  let thickness = 0.03;
  fn bounds() {
    let half = vec3(width * 0.5, height * 0.5, depth * 0.5);
    let r = length(vec3(max(half.x, half.z), 0.0, max(half.x, half.z)));
    return vec3(r + thickness, half.y + thickness, r + thickness);
  }
  fn domain(p) {
    return rotate_y(p, p.y * twist);
  }
  fn distance(p) {
    let half = vec3(width * 0.5, height * 0.5, depth * 0.5);
    return Box.distance(p, half);
  }
  fn distance_post(d, p) {
    return abs(d + sin(p.y * 120.0) * 0.004) - thickness;
  }
};
let statue_mat = Lambert { color: #d9c9a8 };
var statue = TwistedStatue { material: statue_mat };
statue.pos.y = 0.2;
"""]
    names = ["statue"]
    for i in range(n_prims):
        parts.append(_random_material(rng, f"m{i}"))
        parts.append(_random_primitive(rng, f"p{i}", f"m{i}"))
        names.append(f"p{i}")
    parts.append(f"let scene = {balanced(names, '+', 0.0)};\n")
    opts = RenderOptions(width=width, height=height, max_steps=256, max_dist=40.0, epsilon=1.0e-3,
                         step_scale=0.7, camera_z=6.0, fov_y_degrees=45.0)
    settings = RaySettings(max_depth=4, tile_size=64, aa_samples=1)
    return Workload("bricks_4096", "".join(parts), opts, AccelMode.Bricks, settings, 1, "whitted")


def mixed_scene(n_prims: int, seed: int) -> str:
    """One scene of the batched workload: operators over {union, diff, intersect} x
    {round, chamfer, stairs} folded in a balanced tree, Lambert materials."""
    rng = SplitMix64(seed)
    parts = []
    names = []
    for i in range(n_prims):
        col = [rng.uniform(0.2, 0.9) for _ in range(3)]
        parts.append(f"let m{i} = Lambert {{ color: {_vec(*col)} }};\n")
        parts.append(_random_primitive(rng, f"p{i}", f"m{i}"))
        names.append(f"p{i}")
    ops = ["union_round", "union_chamfer", "union_stairs", "diff_round", "diff_chamfer", "diff_stairs",
           "intersect_round", "intersect_chamfer", "intersect_stairs"]

    def tree(ns):
        if len(ns) == 1:
            return ns[0]
        mid = len(ns) // 2
        lhs, rhs = tree(ns[:mid]), tree(ns[mid:])
        op = ops[rng.choice(len(ops))]
        if op.startswith("intersect") or op.startswith("diff"):
            # keep the scene visible: most joins are unions, a quarter subtract/intersect
            if rng.uniform() < 0.75:
                op = "union_" + op.split("_")[1]
        if op.endswith("stairs"):
            return f"{lhs}.{op}({rhs}, 0.1, 3.0)"
        return f"{lhs}.{op}({rhs}, 0.1)"

    parts.append(f"let scene = {tree(names)};\n")
    return "".join(parts)


def log_uniform_counts(n: int, lo: int, hi: int, seed: int) -> List[int]:
    rng = SplitMix64(seed)
    return [int(math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(n)]


LIGHTS_BATCH = ("let key = PointLight { position: vec3(6.0, 9.0, -8.0), color: #fff2d8, intensity: 180.0 };\n"
                "let skylight = EnvLight { color: #cfe0ff, intensity: 0.25 };\n")


@dataclass
class BatchWorkload:
    name: str
    sources: List[str]
    views: List[tuple]          # (scene index, origin, target, fov_y)
    options: RenderOptions
    settings: RaySettings
    accel: AccelMode
    counts: List[int]


def random_views(rng: SplitMix64, n_scenes: int, per_scene: int, radius: float = 16.0) -> List[tuple]:
    """Cameras on a sphere around the scene cube [-5, 5]^3, looking near its centre."""
    views = []
    for sc in range(n_scenes):
        for _ in range(per_scene):
            z = rng.uniform(-0.3, 0.9)          # mostly from above the horizon
            phi = rng.uniform(0.0, 2.0 * math.pi)
            r = math.sqrt(max(1.0 - z * z, 0.0))
            origin = (radius * r * math.cos(phi), radius * z, radius * r * math.sin(phi))
            target = (rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0))
            views.append((sc, origin, target, 45.0))
    return views


def batch_256(n_scenes: int = 256, views_per_scene: int = 4, size: int = 256, lo: int = 8, hi: int = 2048,
              seed: int = 4) -> BatchWorkload:
    """Batched persistent-launch workload: n_scenes seeded scenes with primitive counts
    log-uniform in [lo, hi] (mixed union/diff/intersect x round/chamfer/stairs), each seen
    from views_per_scene random views at size x size: depth preview plus a max_depth-1
    Whitted trace (primary hit + shadowed direct light) with all-Lambert materials."""
    counts = log_uniform_counts(n_scenes, lo, hi, seed)
    sources = [LIGHTS_BATCH + mixed_scene(n, seed * 100003 + k) for k, n in enumerate(counts)]
    views = random_views(SplitMix64(seed + 1), n_scenes, views_per_scene)
    opts = RenderOptions(width=size, height=size, max_steps=128, max_dist=40.0, epsilon=1.0e-3,
                         step_scale=0.7, camera_z=6.0, fov_y_degrees=45.0)
    settings = RaySettings(max_depth=1, tile_size=64, aa_samples=1)
    return BatchWorkload("batch_256", sources, views, opts, settings, AccelMode.Bvh, counts)


@dataclass
class AdjointWorkload:
    name: str
    n_prims: int
    params_true: List[float]
    params_init: List[float]
    views: List[tuple]
    options: RenderOptions
    light_dir: tuple
    normal_step: float
    steps: int
    lr: float


def gauss(rng: SplitMix64) -> float:
    """Standard normal by Box-Muller from two SplitMix64 uniforms."""
    u1 = max(rng.uniform(), 1.0e-12)
    u2 = rng.uniform()
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def adjoint_fit(n_prims: int = 32, n_views: int = 16, size: int = 512, seed: int = 5, sigma: float = 0.05,
                steps: int = 200) -> AdjointWorkload:
    """Differentiable inverse fit: n_prims spheres (centre U[-2,2]^3, radius U[0.4,0.9]) in a
    balanced union_round tree (k U[0.1,0.4]); n_views cameras on a sphere of radius 7;
    target rendered from these parameters, the fit starts from them perturbed by
    N(0, sigma)."""
    rng = SplitMix64(seed)
    th = []
    for _ in range(n_prims):
        th += [rng.uniform(-2.0, 2.0), rng.uniform(-2.0, 2.0), rng.uniform(-2.0, 2.0), rng.uniform(0.4, 0.9)]
    th += [rng.uniform(0.1, 0.4) for _ in range(n_prims - 1)]
    init = [v + sigma * gauss(rng) for v in th]
    views = []
    for k in range(n_views):
        z = rng.uniform(-0.5, 0.9)
        phi = rng.uniform(0.0, 2.0 * math.pi)
        r = math.sqrt(max(1.0 - z * z, 0.0))
        views.append((0, (7.0 * r * math.cos(phi), 7.0 * z, 7.0 * r * math.sin(phi)), (0.0, 0.0, 0.0), 45.0))
    opts = RenderOptions(width=size, height=size, max_steps=256, max_dist=30.0, epsilon=1.0e-5, step_scale=0.7,
                         camera_z=6.0, fov_y_degrees=45.0)
    return AdjointWorkload("adjoint_fit", n_prims, th, init, views, opts, (0.4, 1.0, 0.3), 1.0e-2, steps, 5.0e-3)


WORKLOADS = {"depth_512": depth_512, "whitted_1080p": whitted_1080p, "bricks_4096": bricks_4096,
             "batch_256": batch_256, "adjoint_fit": adjoint_fit}
