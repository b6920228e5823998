"""Host-side pieces of bench.py, on the CPU: the HBM traffic that the roofline
object takes from the committed rocprofv3 PMC summary, and the refusal of
multi-GPU runs."""
import argparse
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _args(subdiv=6, radius=4.0, ladomax=0.25):
    return argparse.Namespace(subdiv=subdiv, radius=radius, ladomax=ladomax)


def _profile_row(kernel):
    """(scaled read MB, scaled write MB) of `kernel` in the committed summary."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_matvecs_sphere122880.md")
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0].strip("`") == kernel:
            return float(cells[4]), float(cells[5])
    raise AssertionError(f"{kernel} not in the profile")


def test_traffic_of_the_roofline_kernel_comes_from_the_profile():
    rd, wr = _profile_row("k_translate_parent")
    t = bench.pmc_traffic("k_translate_parent", _args())
    assert t == round((rd + wr) * 1e6)
    # per launch, averaged over the four levels: above the 40.3 MB
    # algorithmic (the transfer table and shared source rows are read by
    # several XCDs), well below a re-read of every source per child
    assert 4.0e7 < t < 2.0e8
    # the near kernel streams its values once: calibrated counters at the
    # algorithmic bytes (421 MB per launch)
    rd, wr = _profile_row("k_near_box<float2 >")
    assert 3.9e8 < (rd + wr) * 1e6 < 4.5e8


def test_traffic_is_none_for_another_workload_or_kernel():
    assert bench.pmc_traffic("k_translate_parent", _args(subdiv=5)) is None
    assert bench.pmc_traffic("k_translate_parent", _args(ladomax=0.2)) is None
    assert bench.pmc_traffic("k_no_such_kernel", _args()) is None


def test_the_profile_is_calibrated_on_the_copy_kernel():
    # k_copy moves 512 MiB each way per dispatch; the scaled columns must say so
    rd, wr = _profile_row("k_copy")
    assert rd == pytest.approx(536.870912, rel=1e-3)
    assert wr == pytest.approx(536.870912, rel=1e-3)


def test_multi_gpu_runs_are_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]
