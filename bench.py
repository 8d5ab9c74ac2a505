"""Throughput benchmark of the batched HDBSCAN hot path on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload NAME]

A step is one backend call over one batch of B windows (all stages: core
distances, Prim, sort, linkage, condensation, selection, labels) with the
inputs already resident in HBM.  The default batch is the throughput workload
mixture_1024x16: B=256 windows of N=1024 points, D=16, float32, 3-15 Gaussian
blobs + 10 % noise per window, min_cluster_size=5, min_samples=5, epsilon=0.

Protocol (the reference's harness, benchmarks/benchmark.py:20-290, on this
backend):
  1. parity gate: every window's labels against the CPU oracle, label for
     label, and against scikit-learn as partitions on a sample; any mismatch
     exits with status 3 before anything is timed;
  2. the driver's timing: W warmup steps, then exactly K steps between two
     device synchronisations (``value``, ``ms_per_step``);
  3. repetitions: >= 30 single steps, each timed to completion, device
     resident and host to host (host array in, host labels out): median and
     p95 per batch;
  4. the plain exact-distance form (no MFMA Gram filter) measured the same
     way, per-stage HIP-event times, the dominant kernel's roofline, the
     measured VALU issue ceiling, and the CPU baselines (scikit-learn, the
     reference's own comparator, one process per host core; the C oracle).

Prints ONE JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3  # MI355X dense fp32 (vector and matrix), published
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
METRIC = "windows/sec B=256 N=1024 D=16 fp32, label-exact; MFMA-filter vs plain"
EXIT_PARITY = 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# --------------------------------------------------------------------------
# Protocol pieces (CPU-testable: tests/test_bench.py)


def partition_equal(actual, expected):
    """Exact noise mask and a bijection between cluster labels
    (reference benchmarks/benchmark.py:20-33)."""
    actual, expected = np.asarray(actual), np.asarray(expected)
    if actual.shape != expected.shape or not np.array_equal(actual == -1, expected == -1):
        return False
    forward, backward = {}, {}
    for a, e in zip(actual.ravel().tolist(), expected.ravel().tolist()):
        if e == -1:
            continue
        if forward.setdefault(a, e) != e or backward.setdefault(e, a) != a:
            return False
    return True


def timings(operation, warmup, repeats, windows, clock=time.perf_counter):
    """Median / p95 batch time of `repeats` completed calls after `warmup`
    untimed ones (reference benchmarks/benchmark.py:43-57)."""
    for _ in range(warmup):
        operation()
    samples = []
    for _ in range(repeats):
        t0 = clock()
        operation()
        samples.append(clock() - t0)
    median = float(np.median(samples))
    return {"median_batch_ms": median * 1e3, "p95_batch_ms": float(np.quantile(samples, 0.95)) * 1e3,
            "windows_per_second": windows / median, "repeats": repeats, "samples_seconds": samples}


def parity_gate(got, expected, label, exact=True):
    """Counts the windows whose labels equal the expected ones (exact=True:
    label for label; otherwise as partitions).  Exits with EXIT_PARITY on
    any mismatch, before anything is timed."""
    same = [np.array_equal(a, b) if exact else partition_equal(a, b) for a, b in zip(got, expected)]
    if len(got) != len(expected) or not all(same):
        bad = [i for i, s in enumerate(same) if not s]
        log(f"PARITY FAILURE ({label}): {len(bad) + abs(len(got) - len(expected))} of {len(expected)} "
            f"windows differ (first: {bad[:8]})")
        sys.exit(EXIT_PARITY)
    log(f"parity gate ({label}): {len(same)}/{len(expected)} windows "
        f"{'label-exact' if exact else 'partition-equal'}")
    return len(same)


# Keys and types of the JSON line (the driver's contract plus this harness's
# evidence); check_schema raises ValueError on a malformed line.
SCHEMA = {
    "metric": str, "value": float, "unit": str, "n_gpus": int, "steps": int, "warmup": int,
    "ms_per_step": float, "higher_is_better": bool, "scaling": str, "dtype": str, "data": str,
    "config": dict, "label_exact_windows": int, "timing": dict, "roofline": dict, "cpu_baseline": (dict, type(None)),
    "input_sha256": str, "versions": dict,
}
ROOFLINE_KEYS = ("bound", "achieved", "peak", "unit", "frac", "traffic")
CPU_KEYS = ("value", "unit", "cores", "kind", "sample")


def check_schema(line):
    for key, kind in SCHEMA.items():
        if key not in line:
            raise ValueError(f"missing key {key}")
        value = line[key]
        if kind is float and isinstance(value, int) and not isinstance(value, bool):
            value = float(value)
        if not isinstance(value, kind) or (kind is int and isinstance(value, bool)):
            raise ValueError(f"{key} has type {type(value).__name__}")
    if "vs_baseline" not in line:
        raise ValueError("missing key vs_baseline")
    if not line["value"] > 0 or not np.isfinite(line["value"]):
        raise ValueError("value must be positive and finite")
    for key in ROOFLINE_KEYS:
        if key not in line["roofline"]:
            raise ValueError(f"roofline lacks {key}")
    if line["roofline"]["bound"] not in ("hbm", "mfma", "valu"):
        raise ValueError("roofline bound must be hbm, mfma or valu")
    if line["cpu_baseline"] is not None:
        for key in CPU_KEYS:
            if key not in line["cpu_baseline"]:
                raise ValueError(f"cpu_baseline lacks {key}")
        if line["cpu_baseline"]["kind"] not in ("reference", "port"):
            raise ValueError("cpu_baseline kind must be reference or port")
    if "workload" not in line["config"]:
        raise ValueError("config names no workload")
    json.dumps(line, allow_nan=False)  # no NaN / inf in the line
    return line


def input_sha256(x):
    return hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest()


def roofline_entry(bound, achieved, peak, unit, kernel, algorithmic, traffic=None, attainable=None):
    out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
           "traffic": traffic, "kernel": kernel, "algorithmic": algorithmic}
    if attainable:
        out["attainable"] = attainable
        out["frac_attainable"] = achieved / attainable
    return out


def roofline_for(dominant, ms, batch, n, d, precision, valu_tops=None, matrix=False):
    """Roofline of the dominant kernel from its algorithmic work per launch
    and its average duration `ms` (matrix: Prim reads a materialised matrix)."""
    itemsize = 4 if precision == "float32" else 8
    t = ms * 1e-3
    pairs = n * (n - 1) / 2
    if dominant in ("core_gram_kernel", "core_window_kernel"):
        flops = 2.0 * d * n * n * batch
        return roofline_entry("mfma", flops / t / 1e12, BF16_PEAK_TFLOPS, "TFLOP/s", dominant,
                              f"{flops:.4g} FLOP per launch (Gram 2*D*N^2 per window)")
    if dominant in ("prim_spec_kernel", "core_select_kernel") or (dominant == "prim_kernel" and matrix):
        rows = pairs if dominant != "core_select_kernel" else n * n
        nbytes = batch * rows * itemsize
        what = "N(N-1)/2 weights" if dominant != "core_select_kernel" else "N^2 weights"
        return roofline_entry("hbm", nbytes / t / 1e9, HBM_PEAK_GBS, "GB/s", dominant,
                              f"{nbytes:.4g} B per launch ({what} per window)")
    if dominant in ("prim_kernel", "core_topk_kernel", "distance_kernel"):
        # sub, mul, add per coordinate per unordered pair; no FMA can merge them.
        flops = 3.0 * d * pairs * batch
        peak = FP32_PEAK_TFLOPS if precision == "float32" else FP64_PEAK_TFLOPS
        attainable = valu_tops if precision == "float32" else None
        return roofline_entry("valu", flops / t / 1e12, peak, "TFLOP/s", dominant,
                              f"{flops:.4g} FLOP per launch (3*D*N(N-1)/2 per window)", attainable=attainable)
    nbytes = batch * ((n - 1) * (8 + itemsize) + 4 * n)
    return roofline_entry("hbm", nbytes / t / 1e9, HBM_PEAK_GBS, "GB/s", dominant,
                          f"{nbytes} B per launch (edges in + labels out)")


def versions():
    out = {"python": platform.python_version(), "numpy": np.__version__, "platform": platform.platform()}
    try:
        import torch

        out["torch"] = torch.__version__
        out["hip"] = getattr(torch.version, "hip", None)
        if torch.cuda.is_available():
            out["device"] = torch.cuda.get_device_name(0)
    except Exception:  # noqa: BLE001 - informative only
        pass
    try:
        import sklearn

        out["sklearn"] = sklearn.__version__
    except Exception:  # noqa: BLE001
        out["sklearn"] = None
    try:
        from hdbscan_amd import _native

        out["hdbscan_rocm"] = int(_native.lib().hdb_version())
    except Exception:  # noqa: BLE001
        pass
    return out


# --------------------------------------------------------------------------
# Device helpers


def to_device(x, torch):
    pinned = torch.from_numpy(np.ascontiguousarray(x)).pin_memory()
    out = pinned.to("cuda", non_blocking=True)
    torch.cuda.synchronize()
    return out


def to_host(t, torch):
    torch.cuda.synchronize()
    pinned = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    pinned.copy_(t)
    torch.cuda.synchronize()
    return pinned.numpy().copy()


def valu_ceiling(torch):
    """Measured issue rate (operations per second) of the exact distance's
    sub/mul/add mix: 8 workgroups of 256 threads per CU, 24 * iters
    operations per thread."""
    from hdbscan_amd import _native

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    blocks, iters = 8 * cus, 4096
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    ms = ctypes_float()
    _native.check(_native.lib().hdb_probe_valu(None, blocks, iters, ptr(out), ms_ref(ms)))
    ops = blocks * 256 * 24.0 * iters
    return ops / (ms.value * 1e-3) / 1e12, {"blocks": blocks, "iters": iters, "ms": ms.value, "cus": cus}


def ctypes_float():
    import ctypes

    return ctypes.c_float(0.0)


def ms_ref(v):
    import ctypes

    return ctypes.byref(v)


def ptr(t):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())


# --------------------------------------------------------------------------
# CPU baselines


def oracle_labels(x, mcs, ms, eps, precision, threads):
    from oracle import oracle

    def one(w):
        labels, status = oracle.hdbscan(w, mcs, ms, eps, precision=precision)
        assert status == 0
        return labels

    with cf.ThreadPoolExecutor(threads) as pool:
        return np.stack(list(pool.map(one, list(x))))


def cpu_oracle_baseline(x, mcs, ms, eps, precision, threads, budget_s=10.0):
    """C oracle windows/s on the host cores, bounded to about budget_s seconds."""
    from oracle import oracle

    oracle.hdbscan(x[0], mcs, ms, eps, precision=precision)
    t0 = time.perf_counter()
    oracle.hdbscan(x[0], mcs, ms, eps, precision=precision)
    per = time.perf_counter() - t0
    count = int(max(threads, budget_s / max(per, 1e-6) * threads))
    sample = [x[i % len(x)] for i in range(count)]

    def one(w):
        return oracle.hdbscan(w, mcs, ms, eps, precision=precision)[1]

    with cf.ThreadPoolExecutor(threads) as pool:
        list(pool.map(one, sample[:threads]))
        t0 = time.perf_counter()
        list(pool.map(one, sample))
        dt = time.perf_counter() - t0
    return count / dt, count, dt


def _sk_worker_init():
    for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[var] = "1"


def _sk_fit(args):
    w, mcs, ms, eps = args
    from sklearn.cluster import HDBSCAN

    return HDBSCAN(min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps,
                   algorithm="auto", metric="euclidean", cluster_selection_method="eom",
                   allow_single_cluster=False, alpha=1.0, n_jobs=1).fit_predict(w)


def sklearn_baseline(x, mcs, ms, eps, workers, budget_s=10.0):
    """scikit-learn windows/s with one single-threaded spawn process per core
    (reference benchmarks/benchmark.py:90-99, 222-238), start-up excluded,
    bounded to about budget_s seconds of timed work.  Also returns the
    labels of the first `workers` windows (partition check)."""
    import multiprocessing as mp

    _sk_worker_init()
    ctx = mp.get_context("spawn")

    def jobs(count):
        return [(np.asarray(x[i % len(x)]), mcs, ms, eps) for i in range(count)]

    with cf.ProcessPoolExecutor(workers, mp_context=ctx, initializer=_sk_worker_init) as pool:
        first = list(pool.map(_sk_fit, jobs(workers)))  # start-up and import barrier
        t0 = time.perf_counter()
        list(pool.map(_sk_fit, jobs(2 * workers)))
        per = (time.perf_counter() - t0) / (2 * workers)
        count = int(max(workers, budget_s / max(per, 1e-6)))
        work = jobs(count)
        t0 = time.perf_counter()
        list(pool.map(_sk_fit, work, chunksize=max(1, count // (8 * workers))))
        dt = time.perf_counter() - t0
    return count / dt, count, dt, first


# --------------------------------------------------------------------------
# Distance error against float64 (float32 workloads)


def distance_error(xd, x, precision, torch, windows=2):
    """Largest error of the exact float32 distances against a float64
    evaluation of the same points, over the first `windows` windows."""
    from hdbscan_amd import _native

    if precision != "float32":
        return None
    sub = xd[:windows].contiguous()
    b, n, d = (int(s) for s in sub.shape)
    out = torch.empty((b, n, n), dtype=torch.float32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    _native.check(_native.lib().hdb_distances(None, ptr(sub), b, n, d, _native.F32, ptr(out), ptr(status)))
    got = to_host(out, torch).astype(np.float64)
    worst_abs = worst_rel = 0.0
    for w in range(b):
        p = x[w].astype(np.float64)
        sq = (p * p).sum(1)
        ref = np.sqrt(np.maximum(sq[:, None] + sq[None, :] - 2.0 * p @ p.T, 0.0))
        # Re-evaluate the pairs the Gram form cancels badly by differences.
        near = ref < 1e-3 * np.sqrt(sq.max() + 1e-300)
        if near.any():
            i, j = np.nonzero(near)
            ref[i, j] = np.sqrt(((p[i] - p[j]) ** 2).sum(1))
        err = np.abs(got[w] - ref)
        worst_abs = max(worst_abs, float(err.max()))
        mask = ref > 0
        if mask.any():
            worst_rel = max(worst_rel, float((err[mask] / ref[mask]).max()))
    return {"windows": b, "max_abs": worst_abs, "max_rel": worst_rel,
            "unit_roundoff_fp32": 2.0 ** -24}


def bounded_error(xd, x, ms, torch, windows=2):
    """The MFMA-bounded matrix against a float64 evaluation of the same
    points: the share of pairs stored exactly, the exact entries' largest
    relative error, and how far below the float64 distance the certified
    lower bounds sit (relative gap), with the count of bounds above it."""
    from hdbscan_amd import _native

    sub = xd[:windows].contiguous()
    b, n, d = (int(s) for s in sub.shape)
    nbytes = _native.lib().hdb_workspace_bytes(b, n, d, _native.F32, int(ms), _native.FLAG_MATRIX)
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")
    out = torch.empty((b, n, n), dtype=torch.float32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    _native.check(_native.lib().hdb_bounded_distances(None, ptr(sub), b, n, d, _native.F32, int(ms), ptr(out),
                                                      ptr(ws), int(nbytes), ptr(status)))
    got = to_host(out, torch)
    exact_pairs = total = over = 0
    worst_exact = 0.0
    gaps = []
    for w in range(b):
        p = x[w].astype(np.float64)
        sq = (p * p).sum(1)
        ref = np.sqrt(np.maximum(sq[:, None] + sq[None, :] - 2.0 * p @ p.T, 0.0))
        close = ref < 1e-3 * np.sqrt(sq.max() + 1e-300)
        if close.any():
            i, j = np.nonzero(close)
            ref[i, j] = np.sqrt(((p[i] - p[j]) ** 2).sum(1))
        off = ~np.eye(n, dtype=bool)
        m = got[w].astype(np.float64)
        far = np.signbit(got[w]) & off
        near = ~np.signbit(got[w]) & off & (ref > 0)
        exact_pairs += int(near.sum())
        total += int(off.sum())
        if near.any():
            worst_exact = max(worst_exact, float((np.abs(m[near] - ref[near]) / ref[near]).max()))
        if far.any():
            lb = -m[far]
            over += int((lb > ref[far] * (1 + 2.0 ** -22)).sum())
            gaps.append((ref[far] - lb) / np.maximum(ref[far], 1e-300))
    g = np.concatenate(gaps) if gaps else np.zeros(1)
    return {"windows": b, "exact_pair_fraction": exact_pairs / max(total, 1), "exact_max_rel": worst_exact,
            "lower_bound_rel_gap_max": float(g.max()), "lower_bound_rel_gap_median": float(np.median(g)),
            "lower_bounds_above_fp64": over}


# --------------------------------------------------------------------------
# Ragged batches


def bench_ragged(args, spec, torch):
    """Windows of uneven size in one backend call (hdb_cluster_ragged): parity
    gate on every window, then timed steps; prints one JSON line."""
    from hdbscan_amd import _rocm, workloads

    count = args.batch or spec["batch"]
    d, precision = spec["d"], spec["precision"]
    mcs, ms, eps = spec["min_cluster_size"], spec["min_samples"], spec["cluster_selection_epsilon"]
    real = np.float32 if precision == "float32" else np.float64
    windows = [w.astype(real) for w in getattr(workloads, args.workload)(seed=0, count=count)]
    sizes = np.array([len(w) for w in windows])
    log(f"generating {args.workload}: {count} windows, N {sizes.min()}..{sizes.max()} "
        f"(mean {sizes.mean():.0f}), D={d} {precision}")
    packed, offsets = _rocm.pack_windows(windows, precision)
    threads = min(16, host_cores())
    # The autotuner's table for this dimension / precision / min_samples,
    # when one is committed (hdbscan_amd/tuning/).
    from hdbscan_amd import autotune

    tuning = None
    path = args.tuning or autotune.table_path(d, precision, ms)
    if path != "none" and os.path.exists(path):
        autotune.apply(autotune.load(path))
        tuning = os.path.relpath(path, ROOT)
        log(f"tuning table {tuning}")

    def call(**kw):
        return _rocm.cluster_ragged_async(packed, offsets, mcs, ms, eps, precision, flags=args.flags, **kw)

    labels, status, workspace, off_dev = call()
    torch.cuda.synchronize()
    _rocm.raise_status(int(status.item()))
    got = to_host(labels, torch)
    log(f"parity gate: oracle on {count} windows with {threads} threads")
    from oracle import oracle

    def one(w):
        lab, st = oracle.hdbscan(w, mcs, ms, eps, precision=precision)
        assert st == 0
        return lab

    with cf.ThreadPoolExecutor(threads) as pool:
        expected = list(pool.map(one, windows))
    exact = parity_gate([got[offsets[i]:offsets[i + 1]] for i in range(count)], expected, "ragged")
    reuse = dict(out=labels, workspace=workspace, status=status, offsets_dev=off_dev)
    for _ in range(args.warmup):
        call(**reuse)
    status.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call(**reuse)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _rocm.raise_status(int(status.item()))
    value = count * args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed / args.steps * 1e3:.3f} ms/step, {value:.0f} windows/s")

    def resident():
        call(**reuse)
        torch.cuda.synchronize()

    rep = timings(resident, 2, max(30, args.repeats), count)
    log(f"repetitions: median {rep['median_batch_ms']:.3f} ms, p95 {rep['p95_batch_ms']:.3f} ms")
    line = {
        "metric": f"windows/sec ragged {count} windows N log-uniform [64, 4096] D={d} {precision}, label-exact",
        "value": value, "unit": "windows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if precision == "float32" else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "windows": count, "points": int(offsets[-1]), "d": d,
                   "min_cluster_size": mcs, "min_samples": ms, "cluster_selection_epsilon": eps,
                   "flags": args.flags, "tuning": tuning,
                   "launch": "HIP graph replay" if not args.flags & 524288 else "direct launches"},
        "label_exact_windows": exact,
        "timing": {"resident": {k: v for k, v in rep.items() if k != "samples_seconds"}},
        "input_sha256": hashlib.sha256(b"".join(np.ascontiguousarray(w).tobytes() for w in windows)).hexdigest(),
        "versions": versions(),
    }
    print(json.dumps(line, allow_nan=False), flush=True)


# --------------------------------------------------------------------------


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--repeats", type=int, default=30, help="single-step repetitions for median / p95 (>= 30)")
    ap.add_argument("--workload", default="mixture_1024x16")
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--flags", type=int, default=0, help="hdb_cluster tuning flags")
    ap.add_argument("--plain-flags", type=int, default=64,
                    help="flags of the plain exact-distance comparison run (added to --flags)")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-sklearn", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain exact-distance comparison run")
    ap.add_argument("--no-host", action="store_true", help="skip the host-to-host repetitions")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (3 launches) and exit; "
                         "used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    ap.add_argument("--tuning", default=None,
                    help="ragged workloads: tuning table (default: the committed one for the workload, if any; "
                         "'none' for the automatic choice)")
    ap.add_argument("--plant-mismatch", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU, so --gpus {args.gpus} is not run.  Windows are independent: "
              "on several GPUs each rank would take its own share of the batch (one process per GPU, no "
              "exchange between ranks).", file=sys.stderr)
        sys.exit(2)
    if args.repeats < 30:
        ap.error("--repeats must be at least 30")

    import torch

    from hdbscan_amd import _native, _rocm, hdbscan, workloads

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    if args.copy_calibration:
        count = args.copy_calibration * 1024 * 1024 // 4
        src = torch.ones(count, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads and writes count*4 bytes
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": count * 4, "launches": 3}), flush=True)
        return
    spec = dict(workloads.WORKLOADS[args.workload])
    if spec["n"] is None:
        return bench_ragged(args, spec, torch)
    batch = args.batch or spec["batch"]
    n, d, precision = spec["n"], spec["d"], spec["precision"]
    mcs, ms, eps = spec["min_cluster_size"], spec["min_samples"], spec["cluster_selection_epsilon"]
    real = np.float32 if precision == "float32" else np.float64
    gen = getattr(workloads, args.workload)
    log(f"generating {args.workload}: B={batch} N={n} D={d} {precision}")
    x = (gen(seed=0, batch=batch) if spec["batch"] > 1 else gen(seed=0)[None]).astype(real)
    xd = to_device(x, torch)
    threads = min(16, host_cores())

    log(f"parity gate: oracle on {batch} windows with {threads} threads")
    expected = oracle_labels(x, mcs, ms, eps, precision, threads)
    if args.plant_mismatch:
        expected = expected.copy()
        expected[-1, :] = np.where(expected[-1] >= 0, expected[-1] + 1, 0)

    def measure(flags, label):
        """Parity gate on every window, then W warmup and K timed steps,
        single-step repetitions, per-stage HIP-event durations."""
        labels, status, workspace = _rocm.cluster_async(xd, mcs, ms, eps, precision, flags=flags)
        torch.cuda.synchronize()
        _rocm.raise_status(int(status.item()))
        exact = parity_gate(to_host(labels, torch), expected, label)
        out = labels

        def step():
            _rocm.cluster_async(xd, mcs, ms, eps, precision, flags=flags, out=out, workspace=workspace,
                                status=status)

        for _ in range(args.warmup):
            step()
        status.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        _rocm.raise_status(int(status.item()))
        log(f"timed {args.steps} steps ({label}): {elapsed / args.steps * 1e3:.3f} ms/step, "
            f"{batch * args.steps / elapsed:.0f} windows/s")

        def resident():
            step()
            torch.cuda.synchronize()

        rep = timings(resident, 1, args.repeats, batch)
        log(f"repetitions ({label}): median {rep['median_batch_ms']:.3f} ms, p95 {rep['p95_batch_ms']:.3f} ms")
        lib = _native.lib()
        lib.hdb_profile(1)
        stage_sum = {}
        for _ in range(args.profile_steps):
            step()
            torch.cuda.synchronize()
            for name, t in _native.stage_times():
                stage_sum[name] = stage_sum.get(name, 0.0) + t
        lib.hdb_profile(0)
        stages = {k: v / args.profile_steps for k, v in stage_sum.items()}
        log(f"stage ms ({label}): " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))
        return {"value": batch * args.steps / elapsed, "ms_per_step": elapsed / args.steps * 1e3,
                "stage_ms": stages, "exact": exact,
                "resident": {k: v for k, v in rep.items() if k != "samples_seconds"}}

    main_run = measure(args.flags, "default" if not args.flags & 64 else "plain")
    value, ms_per_step, stages, exact = (main_run["value"], main_run["ms_per_step"], main_run["stage_ms"],
                                         main_run["exact"])
    plain = None
    plain_flags = args.flags | args.plain_flags
    if not args.no_plain and plain_flags != args.flags:
        pr = measure(plain_flags, "plain")
        plain = {"value": pr["value"], "ms_per_step": pr["ms_per_step"], "stage_ms": pr["stage_ms"],
                 "flags": plain_flags, "resident": pr["resident"],
                 "mfma_filter_speedup": value / pr["value"]}

    # Host to host: host array in, host labels out, through the backend call
    # (pageable host memory, as a caller's NumPy array would be).
    host_to_host = None
    if not args.no_host:
        def h2h():
            labels = _rocm.cluster(x, mcs, ms, eps, precision, flags=args.flags)
            return labels.cpu().numpy()

        host_to_host = timings(h2h, 2, args.repeats, batch)
        host_to_host.pop("samples_seconds")
        log(f"host to host: median {host_to_host['median_batch_ms']:.3f} ms, "
            f"p95 {host_to_host['p95_batch_ms']:.3f} ms")

    tops, probe = valu_ceiling(torch)
    log(f"VALU issue ceiling (sub/mul/add mix): {tops:.2f} Top/s ({probe})")
    dominant = max(stages, key=stages.get)
    roofline = roofline_for(dominant, stages[dominant], batch, n, d, precision, valu_tops=tops,
                            matrix="distance_kernel" in stages or "gx_matrix_kernel" in stages)
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            traffic = json.load(f).get(f"{args.workload}:{dominant}")
        if traffic is not None:
            roofline["traffic"] = traffic

    # Public API: the batch split into the largest admissible calls (each call
    # reads its status word back, as the reference's API synchronises).
    per_call = max(1, min(batch, (512 * 1024 * 1024) // (2 * n * n * (4 if precision == "float32" else 8))))
    chunks = [xd[i:i + per_call] for i in range(0, batch, per_call)]

    def api():
        for c in chunks:
            hdbscan(c, min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps, precision=precision)
        torch.cuda.synchronize()

    api_rep = timings(api, 1, max(5, args.steps // 4), batch)
    api_rep.pop("samples_seconds")
    log(f"public API ({len(chunks)} x B={per_call}): median {api_rep['median_batch_ms']:.3f} ms")

    dist_err = distance_error(xd, x, precision, torch) if d >= 64 else None
    if dist_err is not None and "gx_matrix_kernel" in stages:
        dist_err["bounded_matrix"] = bounded_error(xd, x, ms, torch)

    cpu = None
    oracle_cpu = None
    sk_partition = None
    if not args.no_cpu:
        cores = threads
        if not args.no_sklearn:
            try:
                import sklearn

                v, count, dt, first = sklearn_baseline(x, mcs, ms, eps, cores)
                sk_partition = sum(partition_equal(a, b) for a, b in zip(expected[:len(first)], first))
                cpu = {"value": v, "unit": "windows/s", "cores": cores, "kind": "reference",
                       "sample": f"{count} windows of {args.workload}, scikit-learn {sklearn.__version__} "
                                 f"HDBSCAN (the reference's CPU comparator, benchmarks/benchmark.py:90-94), "
                                 f"one single-threaded process per core, {dt:.1f} s"}
                log(f"sklearn: {v:.1f} windows/s on {cores} processes; partitions equal on "
                    f"{sk_partition}/{len(first)} sampled windows")
            except Exception as e:  # noqa: BLE001 - baseline is informative only
                log(f"sklearn baseline skipped: {e}")
        v, count, dt = cpu_oracle_baseline(x, mcs, ms, eps, precision, cores)
        oracle_cpu = {"value": v, "unit": "windows/s", "cores": cores, "kind": "port",
                      "sample": f"{count} windows of {args.workload} through the C oracle, {cores} threads, "
                                f"{dt:.1f} s"}
        log(f"cpu oracle: {v:.1f} windows/s on {cores} threads")
        if cpu is None:
            cpu = oracle_cpu

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "windows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if precision == "float32" else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "batch": batch, "n": n, "d": d,
                   "min_cluster_size": mcs, "min_samples": ms, "cluster_selection_epsilon": eps,
                   "path": ("MFMA-bounded matrix" if "gx_matrix_kernel" in stages else
                            "MFMA Gram filter" if any(k.startswith(("gram", "core_gram", "core_window"))
                                                      for k in stages) else "plain exact-distance"),
                   "flags": args.flags},
        "label_exact_windows": exact,
        "timing": {"resident": main_run["resident"], "host_to_host": host_to_host,
                   "public_api": dict(api_rep, calls_per_batch=len(chunks), windows_per_call=per_call)},
        "stage_ms": stages,
        "plain": plain,
        "roofline": roofline,
        "valu_ceiling_tops": tops,
        "cpu_baseline": cpu,
        "oracle_cpu_baseline": oracle_cpu,
        "distance_error_vs_fp64": dist_err,
        "input_sha256": input_sha256(x),
        "versions": versions(),
    }
    if sk_partition is not None:
        # scikit-learn runs with NumPy's default (SIMD) sorting here, whose tie
        # order differs from the scalar order the oracle and the reference
        # reproduce: informative only.
        line["sklearn_default_dispatch_partition_equal"] = {"equal": sk_partition, "sampled": threads}
    if cpu:
        line["gpu_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(check_schema(line)), flush=True)


if __name__ == "__main__":
    main()
