"""bench.py's traffic figure: read from the counters profile only while the profile still
describes the run (same launch shape, frame time within 15%)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _profile_line(bench, prefix):
    with open(bench.TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith(prefix):
                return line.split()
    raise AssertionError(f"no {prefix} line")


def test_profile_traffic_is_used_only_for_the_profiled_run():
    bench = _bench()
    meta = _profile_line(bench, "traffic_launch")
    workload, kernel = meta[1], meta[2]
    kv = dict(x.split("=", 1) for x in meta[3:])
    bw, tw, wpe, un, inl = (int(v) for v in kv["shape"].split(","))
    shape = {"block_waves": bw, "tile": tw, "waves_per_eu": wpe, "march_unroll": un, "shade_inline": inl}
    ms = float(kv["kernel_ms"])
    want = int(_profile_line(bench, "traffic_bytes_per_launch")[-1])

    got, src = bench.measured_traffic(workload, kernel, shape, ms * 1.05)
    assert got == want and "stale" not in src and src["kernel"] == kernel

    got, src = bench.measured_traffic(workload, kernel, dict(shape, block_waves=4 if bw != 4 else 2), ms)
    assert got is None and "launch shape" in src["stale"]
    got, src = bench.measured_traffic(workload, kernel, shape, ms * 1.3)
    assert got is None and "frame" in src["stale"]
    got, src = bench.measured_traffic(workload, "ft_k_no_such_kernel", shape, ms)
    assert got is None and "no counter profile" in src["stale"]
