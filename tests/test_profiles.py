"""The committed rocprofv3 summaries name the kernels bench.py reads them for."""
import importlib.util
import json
import os

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(REPO, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_stage_kernels_have_measured_traffic():
    bench = _bench()
    with open(os.path.join(REPO, "profiles", "alls_bench_hbm_traffic.json")) as f:
        kernels = json.load(f)["kernels"]
    for stage, name in bench.STAGE_KERNEL.items():
        assert name in kernels, f"{stage}: {name} missing from the traffic summary"
        traffic, source = bench.measured_traffic(name)
        assert traffic and traffic > 0 and source


def test_stage_kernels_appear_in_the_kernel_trace():
    bench = _bench()
    text = open(os.path.join(REPO, "profiles", "alls_bench_kernel_stats.md")).read()
    for name in bench.STAGE_KERNEL.values():
        assert f"`{name}`" in text, name
