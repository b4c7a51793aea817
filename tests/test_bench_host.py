"""CPU checks of bench.py's committed inputs: the measured HBM traffic of the chi=1024 chain is
found for the workload it was measured on (the bench's roofline.traffic reads it) and for no
other."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_measured_traffic_matches_only_its_workload():
    b = _bench()
    mix = b.measured_traffic(1024, "mix", "fp32", "L.psi.W12 (mix)")
    r = b.measured_traffic(1024, "mix", "fp32", ".R")
    assert isinstance(mix, int) and isinstance(r, int)
    # the launches move at least their algorithmic bytes (6,725,840 B each)
    assert mix >= 6_725_840 and r >= 6_725_840
    assert b.measured_traffic(1024, "w12", "fp32", "L.psi.W12 (mix)") is None
    assert b.measured_traffic(1024, "mix", "fp64", "L.psi.W12 (mix)") is None
    assert b.measured_traffic(512, "mix", "fp32", "L.psi.W12 (mix)") is None
