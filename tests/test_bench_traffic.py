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


def _profile_lines(bench):
    with open(bench.TRAFFIC_PROFILE) as f:
        return [line.split() for line in f if line.startswith("traffic_bytes_per_launch")]


def test_profile_traffic_is_used_only_for_the_profiled_run():
    bench = _bench()
    lines = _profile_lines(bench)
    assert len(lines) >= 2
    shapes = set()
    for fields in lines:
        workload, kernel = fields[1], fields[2]
        kv = dict(x.split("=", 1) for x in fields[3:5])
        bw, tw, wpe, un, inl = (int(v) for v in kv["shape"].split(","))
        shapes.add((bw, tw, wpe, un, inl))
        shape = {"block_waves": bw, "tile": tw, "waves_per_eu": wpe, "march_unroll": un, "shade_inline": inl}
        ms = float(kv["kernel_ms"])
        want = int(fields[-1])
        # each profiled shape finds its own line
        got, src = bench.measured_traffic(workload, kernel, shape, ms * 1.05)
        assert got == want and "stale" not in src and src["kernel"] == kernel and src["shape"] == kv["shape"]
        got, src = bench.measured_traffic(workload, kernel, shape, ms * 1.3)
        assert got is None and "frame" in src["stale"]
        got, src = bench.measured_traffic(workload, "ft_k_no_such_kernel", shape, ms)
        assert got is None and "no counter profile" in src["stale"]
    # a shape nobody profiled has no traffic figure
    other = {"block_waves": 2, "tile": 32, "waves_per_eu": 1, "march_unroll": 4, "shade_inline": 1}
    assert (2, 32, 1, 4, 1) not in shapes
    got, src = bench.measured_traffic(lines[0][1], lines[0][2], other, 60.0)
    assert got is None and "launch shape" in src["stale"]
