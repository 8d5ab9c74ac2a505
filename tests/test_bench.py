"""The benchmark harness's protocol on the CPU (reference
tests/test_benchmark.py:13-172): the partition truth table, the timing
statistics, the parity gate that stops before any timing, the JSON line's
schema and the refusal of several GPUs."""

import json
import subprocess
import sys

import numpy as np
import pytest

import bench


@pytest.mark.parametrize("actual,expected,equal", [
    ([7, 7, 3, 3, -1], [0, 0, 1, 1, -1], True),
    ([0, 0, 1, 1], [1, 1, 0, 0], True),
    ([-1, -1], [-1, -1], True),
    ([], [], True),
    ([0, 0, 0, 0], [0, 0, 1, 1], False),
    ([0, 1, 2, 2], [0, 0, 1, 1], False),
    ([0, 0, 1, 1], [0, 0, 1, -1], False),
    ([0, -1, 1, 1], [0, 0, 1, -1], False),
    ([0, 1], [0], False),
    ([[0, 0]], [0, 0], False),
])
def test_partition_equal_requires_bijection_and_exact_noise(actual, expected, equal):
    assert bench.partition_equal(actual, expected) is equal
    assert bench.partition_equal(expected, actual) is equal


@pytest.mark.parametrize("warmup", [0, 2])
def test_timings_excludes_warmup_and_reports_median_and_p95(warmup):
    events = []
    ticks = iter([10, 10.25, 20, 20.5, 30, 30.75, 40, 41])

    def clock():
        events.append("clock")
        return next(ticks)

    def operation():
        events.append("operation")

    result = bench.timings(operation, warmup=warmup, repeats=4, windows=3, clock=clock)
    assert events == ["operation"] * warmup + ["clock", "operation", "clock"] * 4
    assert result["samples_seconds"] == [0.25, 0.5, 0.75, 1.0]
    assert result["median_batch_ms"] == 625.0
    assert result["p95_batch_ms"] == pytest.approx(962.5)
    assert result["windows_per_second"] == pytest.approx(4.8)


def test_parity_gate_exits_before_timing_on_a_planted_mismatch(monkeypatch):
    timed = []
    monkeypatch.setattr(bench, "timings", lambda *a, **k: timed.append(1))
    good = np.array([[0, 0, 1, -1], [1, 1, 0, 0]])
    assert bench.parity_gate(good, good.copy(), "ok") == 2
    planted = good.copy()
    planted[1] = [0, 0, 1, 1]  # the same partition, other label numbers
    with pytest.raises(SystemExit) as e:
        bench.parity_gate(planted, good, "planted")
    assert e.value.code == bench.EXIT_PARITY
    assert bench.parity_gate(planted, good, "partition", exact=False) == 2
    planted[0, 3] = 1  # a noise point becomes a member
    with pytest.raises(SystemExit) as e:
        bench.parity_gate(planted, good, "planted", exact=False)
    assert e.value.code == bench.EXIT_PARITY
    with pytest.raises(SystemExit):
        bench.parity_gate(good[:1], good, "missing window")
    assert timed == []


def sample_line():
    return {
        "metric": bench.METRIC, "value": 250000.0, "unit": "windows/s", "n_gpus": 1, "steps": 20, "warmup": 3,
        "ms_per_step": 1.02, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": {"workload": "mixture_1024x16", "batch": 256},
        "label_exact_windows": 256, "timing": {"resident": {"median_batch_ms": 1.0}},
        "roofline": bench.roofline_for("prim_kernel", 0.69, 256, 1024, 16, "float32", valu_tops=40.0),
        "cpu_baseline": {"value": 1700.0, "unit": "windows/s", "cores": 16, "kind": "reference", "sample": "x"},
        "input_sha256": "0" * 64, "versions": {"numpy": np.__version__},
    }


def test_json_line_schema():
    line = bench.check_schema(sample_line())
    assert json.loads(json.dumps(line)) == line
    assert line["roofline"]["bound"] == "valu"
    assert line["roofline"]["frac"] == pytest.approx(3 * 16 * 1024 * 1023 / 2 * 256 / 0.69e-3 / 1e12 / 157.3)
    assert line["roofline"]["frac_attainable"] == pytest.approx(line["roofline"]["achieved"] / 40.0)
    for key in ("metric", "value", "roofline", "cpu_baseline", "input_sha256", "vs_baseline"):
        broken = sample_line()
        del broken[key]
        with pytest.raises(ValueError, match=key):
            bench.check_schema(broken)
    for key, bad in (("value", float("nan")), ("value", -1.0), ("steps", 2.5), ("n_gpus", True)):
        broken = sample_line()
        broken[key] = bad
        with pytest.raises(ValueError):
            bench.check_schema(broken)
    broken = sample_line()
    broken["roofline"]["bound"] = "flops"
    with pytest.raises(ValueError, match="bound"):
        bench.check_schema(broken)
    broken = sample_line()
    broken["cpu_baseline"]["kind"] = "guess"
    with pytest.raises(ValueError, match="kind"):
        bench.check_schema(broken)


def test_roofline_kinds():
    assert bench.roofline_for("prim_kernel", 3.0, 128, 2048, 128, "float32", matrix=True)["bound"] == "hbm"
    assert bench.roofline_for("core_gram_kernel", 0.1, 256, 1024, 16, "float32")["bound"] == "mfma"
    assert bench.roofline_for("core_window_kernel", 0.1, 256, 1024, 16, "float32")["bound"] == "mfma"
    assert bench.roofline_for("tree_kernel", 0.2, 256, 1024, 16, "float32")["bound"] == "hbm"
    r = bench.roofline_for("distance_kernel", 2.5, 128, 2048, 128, "float32", valu_tops=42.0)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s"


@pytest.mark.parametrize("argv,code", [(["--gpus", "2"], 2), (["--repeats", "10"], 2)])
def test_refusals_before_any_device_work(argv, code):
    out = subprocess.run([sys.executable, bench.__file__, *argv], capture_output=True, text=True)
    assert out.returncode == code
    assert out.stdout == ""
