"""`ftc` on the GPU backend: the reference CLI's render entry points (crates/ftc/src/main.rs)
with the same subcommands, flags, defaults, accel choice and exit codes.

  python -m forgedthoughts_amd.ftc [trace flags]          trace (the default, M:270-277)
  python -m forgedthoughts_amd.ftc trace|ray  ...         Whitted PNG (M:48-91, 631-717)
  python -m forgedthoughts_amd.ftc depth|render ...       depth preview PNG (M:92-124, 382-479)
  python -m forgedthoughts_amd.ftc bench ...              Naive / Bvh / Bricks timings (M:125-144, 481-560)
  python -m forgedthoughts_amd.ftc check ...              parse and evaluate only (M:40-47, 352-369)

Exit codes (M:352-560, 743-790): 0 success, 2 no scene given or iterations = 0, 3 the
scene failed to load or evaluate, 4 the render or the PNG failed.  The scene comes from
--scene or FORGEDTHOUGHTS_SCENE (AppConfig::from_env); the accel from --accel, else the
scene's RenderSettings, else Bvh (M:387-390); the options merge the scene's settings under
--width / --height (merged_render_options M:792-829).  --watch re-runs whenever the file's
size or modification time changes, polling every 250 ms (run_with_watch M:743-790).

The progress callback writes the image after each tile as the reference's does, at most
every 250 ms (the last image is always written); the progress line goes to stderr.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from typing import Callable, Optional

from .options import AccelMode, RaySettings, RenderOptions

AOVS = ("depth", "normal", "material-id", "ior", "transmission", "fresnel", "hit-t")


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def resolve_scene_path(scene: Optional[str]) -> Optional[str]:
    """resolve_scene_path with AppConfig::from_env: --scene, else FORGEDTHOUGHTS_SCENE."""
    return scene or os.environ.get("FORGEDTHOUGHTS_SCENE") or None


def default_output_path(scene_path: str) -> str:
    root, ext = os.path.splitext(scene_path)
    return root + ".png" if ext else scene_path + ".png"


def merged_render_options(settings, width: Optional[int], height: Optional[int]) -> RenderOptions:
    """merged_render_options (M:792-829): reference defaults, then the scene's settings,
    then the command line's size."""
    o = RenderOptions()
    for name in ("width", "height", "max_steps", "max_dist", "epsilon", "camera_z", "fov_y_degrees"):
        v = getattr(settings, name)
        if v is not None:
            setattr(o, name, v)
    if settings.step_scale is not None:
        o.step_scale = min(max(float(settings.step_scale), 0.05), 1.0)
    if width is not None:
        o.width = width
    if height is not None:
        o.height = height
    return o


def _load(scene_path: str):
    from .forge.imports import load_scene_file
    return load_scene_file(scene_path)


class _TileWriter:
    """The reference's on_tile: progress and the image so far saved to the output path."""

    def __init__(self, output: str):
        self.output, self.last = output, 0.0

    def __call__(self, progress: dict, image) -> None:
        from .render import save_png
        now = time.perf_counter()
        done, total = progress["tiles_done"], progress["tiles_total"]
        if done == total or now - self.last >= 0.25:
            save_png(image, self.output)
            self.last = now
            _log(f"{done}/{total} tiles {progress['elapsed_ms']} ms")


def _accel(cli: Optional[str], settings) -> AccelMode:
    if cli is not None:
        return AccelMode.parse(cli)
    if settings.accel is not None:
        return AccelMode(settings.accel)
    return AccelMode.Bvh


def run_render_once(scene_path: str, a) -> int:
    from .render import render_preview_progressive_with_accel, save_png
    from .scene import extract_scene_render_settings
    try:
        state = _load(scene_path)
    except Exception as exc:  # parse / evaluate errors
        _log(f"error: {scene_path}: {exc}")
        return 3
    settings = extract_scene_render_settings(state)
    options = merged_render_options(settings, a.width, a.height)
    accel = _accel(a.accel, settings)
    output = a.output or default_output_path(scene_path)
    tile_size = max(a.tile_size, 8)
    start = time.perf_counter()
    try:
        image = render_preview_progressive_with_accel(state, options, accel, tile_size, max(a.aa, 1),
                                                      _TileWriter(output))
        save_png(image, output)
    except Exception as exc:
        _log(f"error: {output}: {exc}")
        return 4
    secs = time.perf_counter() - start
    mpix = options.width * options.height / 1e6
    _log(f"depth preview rendered: {output} {options.width}x{options.height} accel={accel.name} "
         f"tile_size={tile_size} aa={max(a.aa, 1)}")
    _log(f"benchmark: render_secs={secs:.3f}s mpix_per_sec={mpix / max(secs, 1e-12):.3f}")
    return 0


def run_ray_once(scene_path: str, a) -> int:
    from .render import render_ray_progressive_with_accel, save_png
    from .scene import extract_scene_render_settings
    try:
        state = _load(scene_path)
    except Exception as exc:
        _log(f"error: {scene_path}: {exc}")
        return 3
    settings = extract_scene_render_settings(state)
    options = merged_render_options(settings, a.width, a.height)
    accel = _accel(a.accel, settings)
    output = a.output or default_output_path(scene_path)
    tile_size = max(a.tile_size, 8)
    aov = a.debug_aov.replace("-", "_") if a.debug_aov else None
    rs = RaySettings(max_depth=max(a.depth, 1), tile_size=tile_size, aa_samples=max(a.aa, 1), debug_aov=aov)
    start = time.perf_counter()
    try:
        image = render_ray_progressive_with_accel(state, options, accel, rs, _TileWriter(output))
        save_png(image, output)
    except Exception as exc:
        _log(f"error: {output}: {exc}")
        return 4
    secs = time.perf_counter() - start
    mpix = options.width * options.height / 1e6
    _log(f"trace rendered: {output} {options.width}x{options.height} accel={accel.name} depth={rs.max_depth} "
         f"tile_size={tile_size} aa={rs.aa_samples} debug_aov={aov}")
    _log(f"benchmark: render_secs={secs:.3f}s mpix_per_sec={mpix / max(secs, 1e-12):.3f}")
    return 0


def run_check_once(scene_path: str, a) -> int:
    try:
        state = _load(scene_path)
    except Exception as exc:
        _log(f"error: {scene_path}: {exc}")
        return 3
    _log(f"scene parsed and evaluated: {scene_path} bindings={len(state.bindings)}")
    _log("check completed")
    return 0


def run_bench(a) -> dict:
    """run_bench (M:481-560): per accel mode, `warmup` then `iterations` depth PNG renders
    (render_depth_png_with_accel), wall time each; avg / best / worst ms and Mpix/s."""
    from .render import render_depth_png_with_accel
    from .scene import extract_scene_render_settings
    if a.iterations == 0:
        _log("error: iterations must be greater than zero")
        return {"exit": 2}
    scene_path = resolve_scene_path(a.scene)
    if scene_path is None:
        _log("error: missing scene input; pass --scene <path> or set FORGEDTHOUGHTS_SCENE")
        return {"exit": 2}
    try:
        state = _load(scene_path)
    except Exception as exc:
        _log(f"error: {scene_path}: {exc}")
        return {"exit": 3}
    options = merged_render_options(extract_scene_render_settings(state), a.width, a.height)
    _log(f"starting benchmark: {scene_path} {options.width}x{options.height} iterations={a.iterations} "
         f"warmup={a.warmup}")
    import tempfile
    rows = []
    for mode in (AccelMode.Naive, AccelMode.Bvh, AccelMode.Bricks):
        out = os.path.join(tempfile.gettempdir(), f"ftc-bench-{os.getpid()}-{mode.name}.png")
        try:
            for _ in range(a.warmup):
                render_depth_png_with_accel(state, out, options, mode)
            total, best, worst = 0.0, None, 0
            for _ in range(a.iterations):
                t0 = time.perf_counter()
                render_depth_png_with_accel(state, out, options, mode)
                dt = time.perf_counter() - t0
                total += dt
                ms = int(dt * 1000)
                best = ms if best is None else min(best, ms)
                worst = max(worst, ms)
        except Exception as exc:
            _log(f"error: accel={mode.name}: {exc}")
            return {"exit": 4}
        finally:
            if os.path.exists(out):
                os.remove(out)
        avg_ms = total * 1000.0 / a.iterations
        mpix = options.width * options.height / 1e6
        rows.append({"accel": mode.name, "avg_ms": avg_ms, "best_ms": best, "worst_ms": worst,
                     "mpix_per_sec": mpix / (total / a.iterations)})
    _log("benchmark results:")
    for r in rows:
        _log(f"benchmark_result accel={r['accel']} avg_ms={r['avg_ms']:.3f} best_ms={r['best_ms']} "
             f"worst_ms={r['worst_ms']} mpix_per_sec={r['mpix_per_sec']:.3f}")
    return {"exit": 0, "rows": rows}


def file_stamp(path: str):
    try:
        st = os.stat(path)
    except OSError:
        return None
    return st.st_size, int(st.st_mtime * 1000)


def run_with_watch(scene: Optional[str], watch: bool, action: Callable[[str], int],
                   sleep: Callable[[float], None] = time.sleep, max_polls: Optional[int] = None) -> int:
    """run_with_watch (M:743-790); `max_polls` bounds the watch loop (tests)."""
    scene_path = resolve_scene_path(scene)
    if scene_path is None:
        _log("error: missing scene input; pass --scene <path> or set FORGEDTHOUGHTS_SCENE")
        return 2
    if not watch:
        return action(scene_path)
    last = file_stamp(scene_path)
    _log(f"watch mode active: {scene_path}")
    action(scene_path)
    polls = 0
    while max_polls is None or polls < max_polls:
        sleep(0.25)
        polls += 1
        cur = file_stamp(scene_path)
        if cur is not None and cur != last:
            last = cur
            _log(f"change detected, rerunning: {scene_path}")
            action(scene_path)
    return 0


def _render_flags(p, trace: bool):
    p.add_argument("-s", "--scene")
    p.add_argument("-o", "--output")
    p.add_argument("--width", type=int)
    p.add_argument("--height", type=int)
    p.add_argument("--accel", choices=("naive", "bvh", "bricks"))
    if trace:
        p.add_argument("--depth", type=int, default=8)
    p.add_argument("--tile-size", type=int, default=64)
    p.add_argument("--aa", type=int, default=1)
    if trace:
        p.add_argument("--debug-aov", choices=AOVS)
    p.add_argument("--watch", action="store_true")


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="ftc", description="ForgedThoughts scene tools on the MI355X backend. "
                                 "Without a subcommand, ftc runs the trace renderer.")
    ap.add_argument("-v", "--verbose", action="count", default=0)
    sub = ap.add_subparsers(dest="command")
    c = sub.add_parser("check", help="Parse and evaluate a scene without rendering")
    c.add_argument("-s", "--scene")
    c.add_argument("--watch", action="store_true")
    _render_flags(sub.add_parser("trace", aliases=["ray"], help="Render a Whitted-style traced PNG"), True)
    _render_flags(sub.add_parser("depth", aliases=["render"], help="Render a depth preview PNG"), False)
    b = sub.add_parser("bench", help="Benchmark all acceleration backends on the same scene")
    b.add_argument("-s", "--scene")
    b.add_argument("--width", type=int)
    b.add_argument("--height", type=int)
    b.add_argument("--iterations", type=int, default=5)
    b.add_argument("--warmup", type=int, default=1)
    return ap


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    ap = parser()
    known = {"check", "trace", "ray", "depth", "render", "bench", "-h", "--help"}
    if not argv or argv[0] not in known:
        # the default command: trace with the top-level flags (DefaultRayArgs, M:202-239)
        t = argparse.ArgumentParser(prog="ftc")
        t.add_argument("-v", "--verbose", action="count", default=0)
        _render_flags(t, True)
        a = t.parse_args(argv)
        return run_with_watch(a.scene, a.watch, lambda p: run_ray_once(p, a))
    a = ap.parse_args(argv)
    if a.command == "check":
        return run_with_watch(a.scene, a.watch, lambda p: run_check_once(p, a))
    if a.command in ("trace", "ray"):
        return run_with_watch(a.scene, a.watch, lambda p: run_ray_once(p, a))
    if a.command in ("depth", "render"):
        return run_with_watch(a.scene, a.watch, lambda p: run_render_once(p, a))
    return run_bench(a)["exit"]


if __name__ == "__main__":
    sys.exit(main())
