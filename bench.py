"""Throughput benchmark of the batched HDBSCAN hot path on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

A step is one backend call over one batch of B windows (all stages: core
distances, Prim, sort, linkage, condensation, selection, labels) with the
inputs already resident in HBM.  The batch is the throughput workload
mixture_1024x16: B=256 windows of N=1024 points, D=16, float32, 3-15 Gaussian
blobs + 10 % noise per window, min_cluster_size=5, min_samples=5, epsilon=0.
Before timing, every window's labels are checked against the CPU oracle
(label-exact gate); the run aborts on any mismatch.

Prints ONE JSON line.  Extra fields: public-API throughput (the batch split
into the largest calls the 512 MiB admission rule accepts), per-stage kernel
times from HIP events, the roofline of the dominant kernel, the CPU oracle
baseline and scikit-learn on the host cores.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3  # MI355X dense fp32 (vector and matrix)
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


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


def oracle_labels(x, mcs, ms, eps, precision, threads):
    from oracle import oracle

    def one(w):
        labels, status = oracle.hdbscan(w, mcs, ms, eps, precision=precision)
        assert status == 0
        return labels

    with cf.ThreadPoolExecutor(threads) as pool:
        return np.stack(list(pool.map(one, list(x))))


def cpu_oracle_baseline(x, mcs, ms, eps, precision, threads, budget_s=10.0):
    """Oracle windows/s on the host cores, bounded to about budget_s seconds."""
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
    os.environ.setdefault("OMP_NUM_THREADS", "1")


def _sk_fit(args):
    w, mcs, ms, eps = args
    from sklearn.cluster import HDBSCAN

    return HDBSCAN(min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps,
                   algorithm="auto", metric="euclidean", cluster_selection_method="eom",
                   allow_single_cluster=False, alpha=1.0, n_jobs=1).fit_predict(w).max()


def sklearn_baseline(x, mcs, ms, eps, workers, budget_s=10.0):
    """scikit-learn windows/s with one single-threaded process per core,
    bounded to about budget_s seconds of timed work."""
    import multiprocessing as mp

    for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[var] = "1"
    ctx = mp.get_context("spawn")

    def jobs(count):
        return [(x[i % len(x)].astype(np.float64), mcs, ms, eps) for i in range(count)]

    with cf.ProcessPoolExecutor(workers, mp_context=ctx, initializer=_sk_worker_init) as pool:
        list(pool.map(_sk_fit, jobs(workers)))  # start-up and import barrier
        t0 = time.perf_counter()
        list(pool.map(_sk_fit, jobs(4 * workers)))
        per = (time.perf_counter() - t0) / (4 * workers)
        count = int(max(workers, budget_s / max(per, 1e-6)))
        work = jobs(count)
        t0 = time.perf_counter()
        list(pool.map(_sk_fit, work, chunksize=max(1, count // (8 * workers))))
        dt = time.perf_counter() - t0
    return count / dt, count, dt


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
    labels, status, workspace, off_dev = _rocm.cluster_ragged_async(packed, offsets, mcs, ms, eps, precision,
                                                                    flags=args.flags)
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
    exact = sum(int(np.array_equal(got[offsets[i]:offsets[i + 1]], e)) for i, e in enumerate(expected))
    if exact != count:
        log(f"PARITY FAILURE: {count - exact} of {count} windows differ from the oracle")
        sys.exit(3)
    log(f"parity gate: {exact}/{count} windows label-exact")
    for _ in range(args.warmup):
        _rocm.cluster_ragged_async(packed, offsets, mcs, ms, eps, precision, flags=args.flags, out=labels,
                                   workspace=workspace, status=status, offsets_dev=off_dev)
    status.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _rocm.cluster_ragged_async(packed, offsets, mcs, ms, eps, precision, flags=args.flags, out=labels,
                                   workspace=workspace, status=status, offsets_dev=off_dev)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _rocm.raise_status(int(status.item()))
    value = count * args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed / args.steps * 1e3:.3f} ms/step, {value:.0f} windows/s")
    line = {
        "metric": f"windows/sec ragged {count} windows N log-uniform [64, 4096] D={d} {precision}, label-exact",
        "value": value, "unit": "windows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if precision == "float32" else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "windows": count, "points": int(offsets[-1]), "d": d,
                   "min_cluster_size": mcs, "min_samples": ms, "cluster_selection_epsilon": eps,
                   "flags": args.flags},
        "label_exact_windows": exact,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mixture_1024x16")
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--flags", type=int, default=0, help="hdb_cluster tuning flags")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-sklearn", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain exact-distance comparison run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (3 launches) and exit; "
                         "used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU, so --gpus {args.gpus} is not run.  Windows are independent: "
              "on several GPUs each rank would take its own share of the batch (one process per GPU, no "
              "exchange between ranks).", file=sys.stderr)
        sys.exit(2)

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

    def measure(flags, label):
        """Parity gate on every window, then W warmup and K timed steps, then
        per-stage HIP-event durations (events on the launch stream)."""
        labels, status, workspace = _rocm.cluster_async(xd, mcs, ms, eps, precision, flags=flags)
        torch.cuda.synchronize()
        _rocm.raise_status(int(status.item()))
        got = to_host(labels, torch)
        exact = int(sum(np.array_equal(a, b) for a, b in zip(got, expected)))
        if exact != batch:
            log(f"PARITY FAILURE ({label}): {batch - exact} of {batch} windows differ from the oracle")
            sys.exit(3)
        log(f"parity gate ({label}): {exact}/{batch} windows label-exact")
        out = labels
        for _ in range(args.warmup):
            _rocm.cluster_async(xd, mcs, ms, eps, precision, flags=flags, out=out, workspace=workspace,
                                status=status)
        status.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _rocm.cluster_async(xd, mcs, ms, eps, precision, flags=flags, out=out, workspace=workspace,
                                status=status)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        _rocm.raise_status(int(status.item()))
        log(f"timed {args.steps} steps ({label}): {elapsed / args.steps * 1e3:.3f} ms/step, "
            f"{batch * args.steps / elapsed:.0f} windows/s")
        lib = _native.lib()
        lib.hdb_profile(1)
        stage_sum = {}
        for _ in range(args.profile_steps):
            _rocm.cluster_async(xd, mcs, ms, eps, precision, flags=flags, out=out, workspace=workspace,
                                status=status)
            torch.cuda.synchronize()
            for name, t in _native.stage_times():
                stage_sum[name] = stage_sum.get(name, 0.0) + t
        lib.hdb_profile(0)
        stages = {k: v / args.profile_steps for k, v in stage_sum.items()}
        log(f"stage ms ({label}): " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))
        return batch * args.steps / elapsed, elapsed / args.steps * 1e3, stages, exact

    log(f"parity gate: oracle on {batch} windows with {threads} threads")
    expected = oracle_labels(x, mcs, ms, eps, precision, threads)
    value, ms_per_step, stages, exact = measure(args.flags, "default" if not args.flags & 64 else "plain")
    plain = None
    if not args.no_plain and not args.flags & 64:
        pv, pms, pstages, _ = measure(args.flags | 64, "plain")
        plain = {"value": pv, "ms_per_step": pms, "stage_ms": pstages, "flags": args.flags | 64,
                 "mfma_filter_speedup": value / pv}
    dominant = max(stages, key=stages.get)
    pairs = n * (n - 1) / 2
    dist_flops = 3.0 * d * pairs * batch  # sub, mul, add per coordinate per unordered pair
    if dominant == "core_gram_kernel":
        gram_flops = 2.0 * d * n * n * batch
        achieved = gram_flops / (stages[dominant] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": dominant,
                    "algorithmic": f"{gram_flops:.4g} FLOP per launch (Gram 2*D*N^2 per window)"}
    elif dominant == "prim_kernel" and "distance_kernel" in stages:
        # Matrix path: Prim reads each unordered weight once (one row per step).
        itemsize = 4 if precision == "float32" else 8
        algo_bytes = batch * n * (n - 1) / 2 * itemsize
        achieved = algo_bytes / (stages[dominant] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                    "algorithmic": f"{algo_bytes:.4g} B per launch (N(N-1)/2 weights per window)"}
    elif dominant == "core_select_kernel":
        # Selection from the materialised matrix: every row read once.
        itemsize = 4 if precision == "float32" else 8
        algo_bytes = batch * n * n * itemsize
        achieved = algo_bytes / (stages[dominant] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                    "algorithmic": f"{algo_bytes:.4g} B per launch (N^2 weights per window)"}
    elif dominant in ("prim_kernel", "core_topk_kernel", "distance_kernel"):
        peak = FP32_PEAK_TFLOPS if precision == "float32" else FP64_PEAK_TFLOPS
        achieved = dist_flops / (stages[dominant] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": None, "kernel": dominant,
                    "algorithmic": f"{dist_flops:.4g} FLOP per launch (3*D*N(N-1)/2 per window)"}
    else:
        itemsize = 4 if precision == "float32" else 8
        algo_bytes = batch * ((n - 1) * (8 + itemsize) + 4 * n)
        achieved = algo_bytes / (stages[dominant] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                    "algorithmic": f"{algo_bytes} B per launch (edges in + labels out)"}
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
    for c in chunks:
        hdbscan(c, min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    api_steps = max(1, args.steps // 4)
    for _ in range(api_steps):
        for c in chunks:
            hdbscan(c, min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps, precision=precision)
    torch.cuda.synchronize()
    api_value = batch * api_steps / (time.perf_counter() - t0)
    log(f"public API ({len(chunks)} x B={per_call}): {api_value:.0f} windows/s")

    cpu = None
    sk = None
    if not args.no_cpu:
        cores = threads
        v, count, dt = cpu_oracle_baseline(x, mcs, ms, eps, precision, cores)
        cpu = {"value": v, "unit": "windows/s", "cores": cores, "kind": "port",
               "sample": f"{count} windows of {args.workload} through the C oracle, {cores} threads, {dt:.1f} s"}
        log(f"cpu oracle: {v:.1f} windows/s on {cores} threads")
        if not args.no_sklearn:
            try:
                v, count, dt = sklearn_baseline(x, mcs, ms, eps, cores)
                import sklearn

                sk = {"value": v, "unit": "windows/s", "cores": cores,
                      "sample": f"{count} windows, scikit-learn {sklearn.__version__} HDBSCAN, one "
                                f"single-threaded process per core, {dt:.1f} s"}
                log(f"sklearn: {v:.1f} windows/s on {cores} processes")
            except Exception as e:  # noqa: BLE001 - baseline is informative only
                log(f"sklearn baseline skipped: {e}")

    line = {
        "metric": "windows/sec B=256 N=1024 D=16 fp32, label-exact; MFMA-filter vs plain",
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
                   "path": "MFMA Gram filter" if "core_gram_kernel" in stages else "plain exact-distance",
                   "flags": args.flags},
        "label_exact_windows": exact,
        "public_api_windows_per_s": api_value,
        "public_api_calls_per_batch": len(chunks),
        "stage_ms": stages,
        "plain": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "sklearn_baseline": sk,
    }
    if cpu:
        line["gpu_vs_cpu_baseline"] = value / cpu["value"]
    if sk:
        line["gpu_vs_sklearn"] = value / sk["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
