"""The benchmark harness end to end on the GPU, on the smallest workload:
a planted label mismatch stops the run with status 3 before anything is
timed, and a clean run prints one JSON line that passes the schema."""

import json
import subprocess
import sys

import pytest

import bench

pytestmark = pytest.mark.gpu

SMALL = ["--workload", "single_256x8", "--steps", "2", "--warmup", "1", "--no-cpu", "--profile-steps", "1"]


def run(*extra):
    print("bench.py", *SMALL, *extra, flush=True)
    return subprocess.run([sys.executable, bench.__file__, *SMALL, *extra], capture_output=True, text=True,
                          timeout=180)


def test_planted_mismatch_exits_before_timing():
    out = run("--plant-mismatch", "--no-plain", "--no-host")
    assert out.returncode == bench.EXIT_PARITY, out.stderr[-2000:]
    assert "PARITY FAILURE" in out.stderr
    assert "timed" not in out.stderr and "repetitions" not in out.stderr
    assert out.stdout == ""


def test_clean_run_prints_one_valid_line():
    out = run()
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    line = bench.check_schema(json.loads(lines[0]))
    assert line["label_exact_windows"] == 1
    assert line["timing"]["resident"]["repeats"] >= 30
    assert line["timing"]["host_to_host"]["median_batch_ms"] > 0
    assert line["plain"]["mfma_filter_speedup"] > 0
    assert line["valu_ceiling_tops"] > 1
