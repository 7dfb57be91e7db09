"""The benchmark script's command-line contract, checked without a GPU:
more than one GPU is refused with an explanation and a non-zero status
before anything touches a device, the option set parses, and every
workload the bench can run names its configuration and solve method."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=120):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    p = _bench("--gpus", "2")
    assert p.returncode != 0
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_help_lists_the_options():
    p = _bench("--help")
    assert p.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--workload", "--bf16", "--traffic-probe", "--refine", "--top"):
        assert opt in p.stdout, opt


def test_workloads_are_complete():
    sys.path.insert(0, ROOT)
    import bench
    assert bench.WORKLOADS.keys() >= {"chol", "lu", "varp", "q3big"}
    for name, w in bench.WORKLOADS.items():
        assert w["method"] in (1, 3), name
        assert isinstance(w["name"], str) and w["name"], name
        assert len(w["dims"]) in (2, 3), name
    # the roofline traffic figure belongs to the default workload's mixed path
    assert set(bench.GEMM_TRAFFIC) == {("chol", "mixed")}
