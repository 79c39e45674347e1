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
    """Times the oracle on a bounded sample of the same workload on all host cores."""
    import numpy as np

    from oracle import ftref
    opts = workload.options
    threads = os.cpu_count() or 1
    # evenly spaced single rows, as many as fit the budget (first measure one row)
    rows_all = list(range(0, opts.height))
    probe = rows_all[len(rows_all) // 2]
    t0 = time.perf_counter()
    ftref.render_whitted(scene, workload.accel, opts, workload.settings, threads=threads, rows=(probe, probe + 1))
    per_row = max(time.perf_counter() - t0, 1e-3)
    n_rows = max(1, min(opts.height, int(seconds / per_row)))
    step = max(1, opts.height // n_rows)
    picked = rows_all[step // 2::step][:n_rows]
    t0 = time.perf_counter()
    for r in picked:
        ftref.render_whitted(scene, workload.accel, opts, workload.settings, threads=threads, rows=(r, r + 1))
    dt = time.perf_counter() - t0
    rays = len(picked) * opts.width * workload.settings.aa_samples
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{len(picked)} evenly spaced rows x {opts.width} px x {workload.settings.aa_samples} spp "
                      f"of the same frame ({rays} primary rays, {dt:.1f} s), C++ restatement of the "
                      f"reference renderer, {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="whitted_1080p")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from forgedthoughts_amd import synthetic
    from forgedthoughts_amd.forge.evaluator import load_scene_source
    from forgedthoughts_amd.render import GpuScene
    from forgedthoughts_amd.scene import compile_scene

    w = synthetic.WORKLOADS[args.workload]()
    scene = compile_scene(load_scene_source(w.source))
    opts = w.options
    dev = torch.device("cuda:0")
    rgb = torch.empty(opts.height * opts.width * 3, dtype=torch.uint8, device=dev)
    gs = GpuScene(scene, w.accel)

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

    spp = w.settings.aa_samples if w.kind == "whitted" else w.aa
    rays_per_step = opts.width * opts.height * spp
    value = rays_per_step * args.steps / wall / 1e6

    # work accounting: one extra, untimed frame with FLOP counting on
    gs.set_flop_counting(True)
    step()
    gs.synchronize()
    flops = gs.stats().sdf_flops
    gs.set_flop_counting(False)
    st = gs.stats()
    avg_kernel_s = kernel_ms / args.steps / 1e3
    achieved = flops / avg_kernel_s / 1e12 if avg_kernel_s > 0 else 0.0

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
                                   "compile_ms": round(st.compile_ms, 1)}},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "peak_without_packed_fp32": PEAK_FP32_TFLOPS / 2,
                     "kernel_ms": round(kernel_ms / args.steps, 3), "stream_ms": round(stream_ms / args.steps, 3),
                     "sdf_flops_per_launch": int(flops)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, scene, args.cpu_seconds)
    gs.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
