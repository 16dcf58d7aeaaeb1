"""bench.py's host-side contract (no GPU): the one-GPU refusal and the
committed PMC traffic files that the roofline fields are read from."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""


def test_committed_traffic_covers_the_roofline_kernels():
    """The bench's roofline `traffic` comes from profiles/traffic_*.json by
    kernel name: the balanced pair kernel on both systems, the one-launch
    integrator (both step kinds) on the uneven system and the per-width
    packed integrator launches of the dense one."""
    B = _bench()
    for tag in ("uneven1600k_mixed", "dense620k_mixed"):
        t = B.measured_traffic(tag, "k_pair_bal<true, false>")
        assert t is not None and t > 0
    for part, tag in (("k_integrate_all<", "uneven1600k_mixed"),
                      ("k_integrate_packed<", "dense620k_mixed")):
        for suffix in ("false>", "true>"):
            v = B.measured_traffic_sum(tag, part, suffix)
            assert v is not None and v > 0
    assert B.measured_traffic("no_such_workload", "k_pair_bal<true, false>") is None
