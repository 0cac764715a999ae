"""CPU checks of bench.py's contract that need no GPU: the one-GPU exit, the
roofline arithmetic, the traffic figure read from the committed profile, and
the hardware-queue setting made before the GPU is initialised."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(args, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([sys.executable, BENCH] + args, capture_output=True, text=True, env=e, timeout=120)


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_a_message(n):
    p = _run(["--gpus", str(n)])
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_roofline_is_algorithmic_bytes_over_the_phase_time():
    sys.path.insert(0, ROOT)
    import bench
    r = {"group_bytes": 1_600_000_000_000, "group_ms": 250.0, "alg_bytes": 1_640_000_000_000, "total_ms": 400.0}
    rf = bench.roofline(r)
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] == pytest.approx(6400.0)  # 1.6 TB in 0.25 s
    assert rf["frac"] == pytest.approx(0.8)
    assert rf["whole_count"]["achieved"] == pytest.approx(4100.0)
    assert rf["whole_count"]["frac"] == pytest.approx(0.5125)
    # the measured traffic comes from the committed rocprofv3 summary of this workload
    txt = open(bench.PROFILE_TRAFFIC).read()
    assert rf["traffic"] is not None and str(rf["traffic"]) in txt


@pytest.mark.parametrize("given,want", [("4", "8"), (None, "8"), ("16", "16")])
def test_hardware_queues_raised_before_the_gpu_starts(given, want):
    env = {k: v for k, v in os.environ.items() if k != "GPU_MAX_HW_QUEUES"}
    if given is not None:
        env["GPU_MAX_HW_QUEUES"] = given
    code = "import os, sys; sys.path.insert(0, %r); import bench; print(os.environ['GPU_MAX_HW_QUEUES'])" % ROOT
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip() == want
