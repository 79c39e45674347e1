"""Benchmark: Whitted trace at 1920x1080, 16 spp, max depth 4, BVH accel, 64 seeded
primitives in a union_round tree (Lambert / Metal / Dielectric, point + env light, sky).

One step = one full frame rendered into HBM-resident buffers (inputs and outputs stay on
the GPU; no host copies in the timed region).  value = primary rays per second (Mrays/s).
Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak


def cpu_baseline(workload, scene, seconds: float):
    """Times the oracle on a bounded sample of the same workload on all host cores.
    Returns the baseline record and the oracle's image rows it rendered {row: rgb row}."""
    from oracle import ftref
    opts = workload.options
    threads = os.cpu_count() or 1
    if workload.kind == "whitted":
        spp = workload.settings.aa_samples
        render = lambda r: ftref.render_whitted(scene, workload.accel, opts, workload.settings, threads=threads,
                                                rows=(r, r + 1))[0][r]
    else:
        spp = workload.aa
        render = lambda r: ftref.render_depth(scene, workload.accel, opts, workload.aa, threads=threads,
                                              rows=(r, r + 1))[0][r]
    # evenly spaced single rows, as many as fit the budget (first measure one row)
    rows_all = list(range(0, opts.height))
    probe = rows_all[len(rows_all) // 2]
    t0 = time.perf_counter()
    rows = {probe: render(probe)}
    per_row = max(time.perf_counter() - t0, 1e-3)
    n_rows = max(1, min(opts.height, int(seconds / per_row)))
    if n_rows == 1:  # one row already fills the budget: the probe is the sample
        picked, dt = [probe], per_row
    else:
        step = max(1, opts.height // n_rows)
        picked = rows_all[step // 2::step][:n_rows]
        t0 = time.perf_counter()
        for r in picked:
            rows[r] = render(r)
        dt = time.perf_counter() - t0
    rays = len(picked) * opts.width * spp
    return ({"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
             "sample": f"{len(picked)} evenly spaced rows x {opts.width} px x {spp} spp "
                       f"of the same frame ({rays} primary rays, {dt:.1f} s), C++ restatement of the "
                       f"reference renderer, {threads} threads"}, rows)


def frame_parity(workload, scene, frame, oracle_rows, f64_rows: int):
    """The timed frame's rows against the oracle's rows of the same frame (bytes within 1 of
    255), and `f64_rows` of them against the oracle in FP64 (the error of the f32 frame
    against an f64 march and shading of the same algorithm)."""
    import numpy as np

    from oracle import ftref
    opts = workload.options
    threads = os.cpu_count() or 1
    rows = sorted(oracle_rows)
    g = frame[rows].astype(int)
    c = np.stack([oracle_rows[r] for r in rows]).astype(int)
    d = np.abs(g - c).max(axis=-1)
    out = {"rows": rows, "pixels": int(d.size), "frac_within_1": round(float((d <= 1).mean()), 6),
           "max_abs_diff": int(d.max()), "mean_abs_diff": round(float(np.abs(g - c).mean()), 5)}
    if f64_rows > 0:
        pick = rows[len(rows) // 2:len(rows) // 2 + f64_rows]
        f64 = []
        for r in pick:
            if workload.kind == "whitted":
                img = ftref.render_whitted(scene, workload.accel, opts, workload.settings, threads=threads,
                                           precision=1, rows=(r, r + 1))[0]
            else:
                img = ftref.render_depth(scene, workload.accel, opts, workload.aa, threads=threads, precision=1,
                                         rows=(r, r + 1))[0]
            f64.append(img[r])
        c64 = np.stack(f64).astype(int)
        d64 = np.abs(frame[pick].astype(int) - c64).max(axis=-1)
        d3264 = np.abs(np.stack([oracle_rows[r] for r in pick]).astype(int) - c64).max(axis=-1)
        out["vs_fp64"] = {"rows": pick, "gpu_frac_within_1": round(float((d64 <= 1).mean()), 6),
                          "gpu_max_abs_diff": int(d64.max()),
                          "oracle_f32_frac_within_1": round(float((d3264 <= 1).mean()), 6),
                          "oracle_f32_max_abs_diff": int(d3264.max())}
    return out


def cpu_baseline_batch(w, scenes, seconds: float):
    """Oracle time per scene on a bounded sample: a few views spread over the primitive-count
    range, a few rows of each (depth + Whitted), scaled to whole scenes."""
    import numpy as np

    from oracle import ftref
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import with_camera
    threads = os.cpu_count() or 1
    opts, rs = w.options, w.settings
    by_cost = sorted(range(len(w.views)), key=lambda v: w.counts[w.views[v][0]])
    picks = [by_cost[int((k + 0.5) * len(by_cost) / 8)] for k in range(8)]
    rows_per_view = 2
    spent, per_view = 0.0, []
    for v in picks:
        sc, o, t, fov = w.views[v]
        ref = with_camera(scenes[sc], o, t, fov)
        step = max(1, opts.height // rows_per_view)
        t0 = time.perf_counter()
        for r in range(step // 2, opts.height, step):
            ftref.render_depth(ref, w.accel, opts, 1, threads=threads, rows=(r, r + 1))
            ftref.render_whitted(ref, w.accel, opts, rs, threads=threads, rows=(r, r + 1))
        dt = time.perf_counter() - t0
        spent += dt
        per_view.append(dt / rows_per_view * opts.height)
        if spent > seconds:
            break
    view_s = float(np.mean(per_view))
    per_scene = view_s * len(w.views) / len(scenes)
    return {"value": 1.0 / per_scene, "unit": "scenes/s", "cores": threads, "kind": "port",
            "sample": f"{len(per_view)} views spread over the primitive-count range, {rows_per_view} rows each "
                      f"(depth + Whitted), {spent:.1f} s, scaled to whole views; C++ restatement of the "
                      f"reference renderer, {threads} threads"}


def bench_batch(args, w):
    import torch

    from forgedthoughts_amd import abi
    from forgedthoughts_amd.forge.evaluator import load_scene_source
    from forgedthoughts_amd.render import GpuBatch
    from forgedthoughts_amd.scene import compile_scene

    t0 = time.perf_counter()
    scenes = [compile_scene(load_scene_source(src)) for src in w.sources]
    print(f"[bench] {len(scenes)} scenes compiled in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    opts, rs = w.options, w.settings
    n = len(w.views)
    per = opts.width * opts.height * 3
    dev = torch.device("cuda:0")
    dd = torch.empty(n * per, dtype=torch.uint8, device=dev)
    dw = torch.empty(n * per, dtype=torch.uint8, device=dev)
    gb = GpuBatch(scenes, w.accel)
    views = gb.views(w.views)
    both = abi.BATCH_DEPTH | abi.BATCH_WHITTED

    def run(flags, steps, label):
        # a 256-scene step takes tens of seconds (several minutes with one launch per view):
        # every warm-up and timed step reports when it ends
        for k in range(args.warmup):
            t1 = time.perf_counter()
            gb.render_device(views, opts, rs, flags, dd.data_ptr(), dw.data_ptr())
            gb.synchronize()
            print(f"[bench] {label} warm-up {k + 1}/{args.warmup}: {(time.perf_counter() - t1) * 1e3:.1f} ms",
                  file=sys.stderr, flush=True)
        ms, t1 = 0.0, time.perf_counter()
        for k in range(steps):
            gb.render_device(views, opts, rs, flags, dd.data_ptr(), dw.data_ptr())
            gb.synchronize()
            ms += gb.stats().last_kernel_ms
            print(f"[bench] {label} step {k + 1}/{steps} done", file=sys.stderr, flush=True)
        return (time.perf_counter() - t1) / steps, ms / steps

    wall_b, kms_b = run(both | abi.BATCH_GRAPH, args.steps, "batched")
    print(f"[bench] batched: {wall_b * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    wall_s, kms_s = run(both | abi.BATCH_SEPARATE, max(1, args.steps // 2), "separate")
    print(f"[bench] separate: {wall_s * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    gb.set_flop_counting(True)
    t1 = time.perf_counter()
    gb.render_device(views, opts, rs, both, dd.data_ptr(), dw.data_ptr())
    gb.synchronize()
    flops = gb.stats().sdf_flops
    print(f"[bench] FLOP-counting frame: {(time.perf_counter() - t1) * 1e3:.1f} ms", file=sys.stderr, flush=True)
    gb.set_flop_counting(False)
    achieved = flops / (kms_b / 1e3) / 1e12
    line = {
        "metric": "batched scenes/s",
        "value": round(len(scenes) / wall_b, 4),
        "unit": "scenes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall_b * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "scenes": len(scenes), "views": n, "width": opts.width,
                   "height": opts.height, "prims": [min(w.counts), max(w.counts), sum(w.counts)],
                   "max_depth": rs.max_depth, "images": "depth + whitted per view", "launch": "batched, hip graph"},
        "separate_launches": {"value": round(len(scenes) / wall_s, 4), "unit": "scenes/s",
                              "ms_per_step": round(wall_s * 1e3, 3), "kernel_launches": 1 + 3 * n,
                              "speedup_of_batched": round(wall_s / wall_b, 3)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "peak_without_packed_fp32": PEAK_FP32_TFLOPS / 2, "kernel_ms": round(kms_b, 3),
                     "sdf_flops_per_launch": int(flops)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_batch(w, scenes, args.cpu_seconds)
    gb.close()
    print(json.dumps(line), flush=True)


def bench_adjoint(args, w):
    import numpy as np

    from forgedthoughts_amd.diff import DiffRenderer

    dr = DiffRenderer(w.n_prims, w.views, w.options, w.light_dir, w.normal_step)
    dr.set_params(w.params_true)
    tt, st = dr.forward()
    dr.set_params(w.params_init)
    dr.set_target(tt, st)
    warm = max(args.warmup, 1)
    dr.fit(warm, w.lr)                      # warm-up steps (part of the same fit)
    steps = w.steps - warm
    t0 = time.perf_counter()
    losses = dr.fit(steps, w.lr)
    wall = time.perf_counter() - t0
    kernel_ms = dr.stats().last_kernel_ms
    p = dr.get_params()
    err0 = float(np.abs(np.float32(w.params_init) - np.float32(w.params_true)).mean())
    err1 = float(np.abs(p - np.float32(w.params_true)).mean())
    npix = len(w.views) * w.options.width * w.options.height
    # work accounting: one untimed step (forward + loss + adjoint) on the FLOP-counting kernels
    dr.set_flop_counting(True)
    dr.loss_grad()
    step_flops = dr.stats().sdf_flops
    dr.set_flop_counting(False)
    step_s = kernel_ms / steps / 1e3
    achieved = step_flops / step_s / 1e12 if step_s > 0 else 0.0
    line = {
        "metric": "adjoint grad-steps/s",
        "value": round(steps / wall, 4),
        "unit": "grad-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": round(wall / steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (fp64 reduction)",
        "data": "synthetic",
        "config": {"workload": w.name, "prims": w.n_prims, "params": 5 * w.n_prims - 1, "views": len(w.views),
                   "width": w.options.width, "height": w.options.height, "optimizer": f"adam lr {w.lr}",
                   "step": "forward + L2 loss + adjoint + fp64 reduction + adam, on the device"},
        "fit": {"loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                "mean_abs_param_error": [round(err0, 5), round(err1, 5)]},
        "kernel_ms_per_step": round(kernel_ms / steps, 3),
        "pixels_per_step": npix,
        # the step's tree-pass FLOPs (forward march + adjoint reverse passes, counted per
        # lane on the device) over the step's kernel time (loss, reduction and Adam included)
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "peak_without_packed_fp32": PEAK_FP32_TFLOPS / 2, "flops_per_step": int(step_flops),
                     "kernel_ms_per_step": round(kernel_ms / steps, 3)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_adjoint(w, args.cpu_seconds)
    dr.close()
    print(json.dumps(line), flush=True)


def cpu_baseline_adjoint(w, seconds: float):
    """The oracle's f32 forward of every view at a reduced resolution, scaled by the pixel
    count to a full step: an upper bound on CPU grad-steps/s (a CPU gradient step costs
    more than one forward)."""
    import copy

    from oracle import ftref
    threads = os.cpu_count() or 1
    size = 128
    o = copy.copy(w.options)
    o.width = o.height = size
    t0 = time.perf_counter()
    n_done = 0
    for v in w.views:
        ftref.diff_forward(w.n_prims, w.params_true, [v], o, w.light_dir, w.normal_step, precision=0,
                           threads=threads)
        n_done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    scale = (w.options.width * w.options.height) / (size * size) * len(w.views) / n_done
    return {"value": 1.0 / (dt * scale), "unit": "grad-steps/s", "cores": threads, "kind": "port",
            "sample": f"forward only (a lower bound on the cost of a CPU gradient step): {n_done} views at "
                      f"{size}x{size} in {dt:.1f} s, scaled by pixel count to {len(w.views)} views at "
                      f"{w.options.width}x{w.options.height}; {threads} threads"}


def copy_calibration(mib: int):
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n, "copy_calibration_bytes_written": n, "launches": 3}),
          flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "whitted_1080p_counters.md")


def measured_traffic(workload: str, kernel: str, shape: dict, kernel_ms: float):
    """HBM bytes per launch of `kernel` on `workload` as measured with the FETCH_SIZE /
    WRITE_SIZE counters (scaled by a plain copy of known size) and recorded in the counters
    profile, one line per profiled launch shape
    (`traffic_bytes_per_launch WORKLOAD KERNEL shape=BW,TILE,WPE,UNROLL,INLINE kernel_ms=MS = BYTES`),
    with where it came from; (None, reason) when the profile holds no line for this workload,
    kernel and launch shape, or the profiled frame took more than 15% longer or shorter than
    this run's, so the kernel has changed since and its traffic must be measured again."""
    entries = []
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                fields = line.split()
                if len(fields) < 4 or fields[0] != "traffic_bytes_per_launch" or fields[1:3] != [workload, kernel]:
                    continue
                meta = dict(kv.split("=", 1) for kv in fields[3:] if "=" in kv and not kv.startswith("="))
                entries.append((meta, int(line.split("=")[-1].strip())))
    except OSError:
        pass
    if not entries:
        return None, {"stale": f"no counter profile for {workload} / {kernel}"}
    src = {"profile": os.path.relpath(TRAFFIC_PROFILE, ROOT), "kernel": kernel,
           "counters": "FETCH_SIZE, WRITE_SIZE of one frame, each scaled by a 1 GiB copy"}
    want = ",".join(str(shape[k]) for k in ("block_waves", "tile", "waves_per_eu", "march_unroll", "shade_inline"))
    match = [(m, b) for m, b in entries if m.get("shape") == want]
    if not match:
        return None, dict(src, stale=f"profiled launch shapes {[m.get('shape') for m, _ in entries]}, this run {want}")
    meta, bytes_ = match[0]
    src["shape"] = want
    prof_ms = float(meta.get("kernel_ms", "0"))
    if not (0.85 * prof_ms <= kernel_ms <= 1.15 * prof_ms):
        return None, dict(src, stale=f"profiled frame {prof_ms} ms, this run {kernel_ms:.3f} ms")
    return bytes_, dict(src, profiled_kernel_ms=prof_ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="whitted_1080p")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (no autotuning)")
    ap.add_argument("--launch", default="", metavar="BW,TILE,WPE,UNROLL,INLINE",
                    help="single-scene workloads: use this launch shape instead of autotuning (block waves, "
                         "tile width (0: sample-major), waves per EU, march unroll, shading inlined)")
    ap.add_argument("--fp64-rows", type=int, default=1,
                    help="rows of the timed frame also checked against the oracle in FP64")
    ap.add_argument("--scenes", type=int, default=0, help="batch workloads: number of scenes (default: the workload's)")
    ap.add_argument("--atlas", type=int, default=0, choices=(0, 16, 32),
                    help="single-scene workloads: coarse march steps on an FP16 / FP32 distance atlas")
    ap.add_argument("--atlas-voxel", type=float, default=0.02)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB, three times (the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from forgedthoughts_amd import synthetic
    from forgedthoughts_amd.forge.evaluator import load_scene_source
    from forgedthoughts_amd.options import RaySettings
    from forgedthoughts_amd.render import GpuScene
    from forgedthoughts_amd.scene import compile_scene

    w = synthetic.WORKLOADS[args.workload](n_scenes=args.scenes) if args.scenes else synthetic.WORKLOADS[args.workload]()
    if isinstance(w, synthetic.BatchWorkload):
        return bench_batch(args, w)
    if isinstance(w, synthetic.AdjointWorkload):
        return bench_adjoint(args, w)
    scene = compile_scene(load_scene_source(w.source))
    opts = w.options
    dev = torch.device("cuda:0")
    rgb = torch.empty(opts.height * opts.width * 3, dtype=torch.uint8, device=dev)
    gs = GpuScene(scene, w.accel)
    if args.atlas:
        gs.build_atlas(args.atlas_voxel, args.atlas)
    # launch shape: the autotuner's choice for this scene and request (outside the timed region)
    from forgedthoughts_amd import abi
    kind = abi.KIND_WHITTED if w.kind == "whitted" else abi.KIND_DEPTH
    tune = None
    if args.launch:
        shape = [int(v) for v in args.launch.split(",")]
        if len(shape) != 5:
            print("--launch takes five comma-separated integers", file=sys.stderr)
            sys.exit(2)
        gs.set_launch_config(kind, *shape)
        tune = {"chosen": gs.launch_config(kind).as_dict(), "tuned": False, "forced": True}
    elif not args.no_tune:
        t_tune = time.perf_counter()
        rep = gs.autotune(kind, opts, w.settings if w.kind == "whitted" else RaySettings(1, 64, w.aa))
        tune = {"chosen": rep.best.as_dict(), "tuned": True, "cached": bool(rep.cached),
                "candidates": rep.candidates, "tune_render_spp": rep.tune_aa_samples,
                "chosen_ms": round(rep.best_ms, 3), "default_ms": round(rep.default_ms, 3),
                "speedup_vs_default": round(rep.default_ms / rep.best_ms, 4) if rep.best_ms > 0 else None,
                "search_s": round(time.perf_counter() - t_tune, 1)}
        print(f"[bench] autotune {tune}", file=sys.stderr, flush=True)

    def step():
        if w.kind == "whitted":
            gs.render_whitted_device(opts, w.settings, rgb.data_ptr())
        else:
            gs.render_depth_device(opts, w.aa, rgb.data_ptr())

    for _ in range(args.warmup):
        step()
        gs.synchronize()

    stream = torch.cuda.ExternalStream(gs.stream(), device=dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    kernel_ms = 0.0
    for _ in range(args.steps):
        step()
        gs.synchronize()
        kernel_ms += gs.stats().last_kernel_ms
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stream_ms = ev0.elapsed_time(ev1)
    frame = rgb.cpu().numpy().reshape(opts.height, opts.width, 3)  # the last timed frame

    spp = w.settings.aa_samples if w.kind == "whitted" else w.aa
    rays_per_step = opts.width * opts.height * spp
    value = rays_per_step * args.steps / wall / 1e6

    # the depth preview of the same scene and image size (1 spp), timed the same way
    depth = None
    if w.kind == "whitted":
        for _ in range(args.warmup):
            gs.render_depth_device(opts, 1, rgb.data_ptr())
            gs.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            gs.render_depth_device(opts, 1, rgb.data_ptr())
            gs.synchronize()
        dwall = (time.perf_counter() - t1) / args.steps
        depth = {"value": round(opts.width * opts.height / dwall / 1e6, 4), "unit": "Mrays/s",
                 "ms_per_step": round(dwall * 1e3, 3), "spp": 1}

    # work accounting: one extra, untimed frame with FLOP counting on
    gs.set_flop_counting(True)
    step()
    gs.synchronize()
    flops = gs.stats().sdf_flops
    gs.set_flop_counting(False)
    st = gs.stats()
    avg_kernel_s = kernel_ms / args.steps / 1e3
    achieved = flops / avg_kernel_s / 1e12 if avg_kernel_s > 0 else 0.0
    # the dominant kernel actually launched (the Whitted kernel's ray-stack size follows the
    # scene's materials), as named in kernel traces and in the counters profile
    traffic_kernel = f"ft_k_whitted_{st.ray_stack}" if w.kind == "whitted" else "ft_k_depth"
    cfg = gs.launch_config(kind)
    traffic, traffic_src = measured_traffic(
        w.name, traffic_kernel,
        {"block_waves": cfg.block_waves, "tile": cfg.tile_width, "waves_per_eu": cfg.waves_per_eu,
         "march_unroll": cfg.march_unroll, "shade_inline": cfg.shade_inline}, kernel_ms / args.steps)

    line = {
        "metric": "Mrays/s depth+Whitted 1080p",
        "value": round(value, 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "width": opts.width, "height": opts.height, "spp": spp,
                   "max_depth": w.settings.max_depth if w.settings else None, "accel": w.accel.name.lower(),
                   "max_steps": opts.max_steps, "epsilon": opts.epsilon, "nodes": int(len(scene.nodes)),
                   "specialised": {"scene": bool(st.flags & 1), "functions": bool(st.flags & 2),
                                   "compile_ms": round(st.compile_ms, 1)},
                   "launch": tune if tune is not None else {"chosen": gs.launch_config(kind).as_dict(),
                                                            "tuned": False}},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "peak_without_packed_fp32": PEAK_FP32_TFLOPS / 2,
                     "kernel_ms": round(kernel_ms / args.steps, 3), "stream_ms": round(stream_ms / args.steps, 3),
                     "sdf_flops_per_launch": int(flops)},
    }
    if depth is not None:
        line["depth_preview"] = depth
    if args.atlas:
        # coarse steps read 8 brick samples and one cell entry each (no cache credit)
        per_sample = 8 * (2 if args.atlas == 16 else 4) + 4
        atlas_bytes = st.atlas_samples * per_sample
        gbs = atlas_bytes / avg_kernel_s / 1e9 if avg_kernel_s > 0 else 0.0
        line["config"]["atlas"] = {"bits": args.atlas, "voxel": args.atlas_voxel, "bricks": int(st.atlas_bricks),
                                   "bytes": int(st.atlas_bytes)}
        line["atlas_roofline"] = {"bound": "hbm", "achieved": round(gbs, 2), "peak": 8000.0, "unit": "GB/s",
                                  "frac": round(gbs / 8000.0, 5), "coarse_samples": int(st.atlas_samples),
                                  "bytes_per_sample": per_sample}
    if not args.no_cpu:
        line["cpu_baseline"], oracle_rows = cpu_baseline(w, scene, args.cpu_seconds)
        line["parity"] = frame_parity(w, scene, frame, oracle_rows, args.fp64_rows)
    gs.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
