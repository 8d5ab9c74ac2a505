"""Autotuner of ragged batches (hdb_cluster_ragged).

For every size bucket of the ragged entry (windows of up to 64, 128, 256,
512, 1024, 2048 and 4096 points) it measures launch configurations on
synthetic windows of that bucket and keeps the fastest:

  vertices_per_thread  Prim vertices per thread, i.e. the workgroup width
                       (flag bits 8..11: automatic, 1, 2, 4)
  gram_row_sets        16-row sets per wave of the MFMA Gram candidate kernel,
                       its tile height (bits 12..13: automatic, 1, 2, 4)
  repack               Prim repacking interval (bits 14..15: 128, never, 32, 64)
  arena                LDS or device memory for the tree arena and the Prim
                       points (HDB_FLAG_GLOBAL_ARENA, HDB_FLAG_NO_POINT_LDS)

one dimension at a time (coordinate descent from the automatic choice).
Every candidate must give the labels of the automatic configuration, window
for window; tuning never changes results.  The table (JSON) keeps every
measurement next to the chosen flags; ``apply`` installs it with
hdb_ragged_set_tuning for later ragged calls of the process.

    python -m hdbscan_amd.autotune --d 16 --precision float32 --min-samples 5 \
        --out hdbscan_amd/tuning/ragged_float32_d16_ms5.json
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _native

BUCKETS = (64, 128, 256, 512, 1024, 2048, 4096)
DIMENSIONS = (
    ("vertices_per_thread", (0, 1 << 8, 2 << 8, 4 << 8)),
    ("gram_row_sets", (0, 1 << 12, 2 << 12, 3 << 12)),
    ("repack", (0, 1 << 14, 2 << 14, 3 << 14)),
    ("arena", (0, _native.FLAG_GLOBAL_ARENA, _native.FLAG_NO_POINT_LDS,
               _native.FLAG_GLOBAL_ARENA | _native.FLAG_NO_POINT_LDS)),
)
TUNE_MASK = 0xFF00 | _native.FLAG_GLOBAL_ARENA | _native.FLAG_NO_POINT_LDS
HERE = os.path.dirname(os.path.abspath(__file__))


def table_path(d, precision, min_samples):
    return os.path.join(HERE, "tuning", f"ragged_{precision}_d{d}_ms{min_samples}.json")


def describe(flags):
    vpt = (flags >> 8) & 15
    rsets = (flags >> 12) & 3
    repack = (flags >> 14) & 3
    return {"vertices_per_thread": vpt or "auto", "gram_row_sets": {0: "auto", 1: 1, 2: 2, 3: 4}[rsets],
            "repack": {0: 128, 1: "never", 2: 32, 3: 64}[repack],
            "tree_arena": "device" if flags & _native.FLAG_GLOBAL_ARENA else "lds",
            "prim_points": "device" if flags & _native.FLAG_NO_POINT_LDS else "lds"}


def bucket_windows(top, d, count, seed):
    """`count` Gaussian-mixture windows (3-15 blobs + 10 % noise) with sizes
    uniform in the bucket (top / 2, top] (2..64 for the first bucket)."""
    rng = np.random.default_rng(seed)
    lo = 2 if top == BUCKETS[0] else top // 2 + 1
    out = []
    for n in rng.integers(lo, top + 1, count):
        k = int(rng.integers(3, 16))
        centres = rng.uniform(-10, 10, (k, d))
        owner = rng.integers(0, k, n)
        x = centres[owner] + rng.normal(size=(n, d)) * rng.uniform(0.5, 1.5, k)[owner, None]
        noise = int(n) // 10
        if noise:
            x[:noise] = rng.uniform(x.min(0), x.max(0), (noise, d))
        out.append(x)
    return out


class _Call:
    """One ragged problem with fixed device buffers."""

    def __init__(self, windows, precision, mcs, ms, eps):
        import torch

        self.torch = torch
        self.lib = _native.lib()
        real = np.float32 if precision == "float32" else np.float64
        self.offsets = np.zeros(len(windows) + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum([len(w) for w in windows])
        host = np.ascontiguousarray(np.concatenate(windows), dtype=real)
        self.x = torch.from_numpy(host).pin_memory().to("cuda")
        self.off_dev = torch.from_numpy(self.offsets).to("cuda")
        self.labels = torch.empty(int(self.offsets[-1]), dtype=torch.int32, device="cuda")
        self.status = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.code = _native.F32 if precision == "float32" else _native.F64
        self.args = (len(windows), int(host.shape[1]), int(mcs), int(ms), float(eps))
        self.ws = None

    def run(self, flags):
        w, d, mcs, ms, eps = self.args
        off = self.offsets.ctypes.data_as(ctypes.c_void_p)
        need = self.lib.hdb_ragged_workspace_bytes(off, w, d, self.code, ms, flags)
        if need == 0:
            return False
        if self.ws is None or self.ws.numel() < need:
            self.ws = self.torch.empty(need, dtype=self.torch.uint8, device="cuda")
        stream = ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        _native.check(self.lib.hdb_cluster_ragged(
            stream, ctypes.c_void_p(self.x.data_ptr()), off, ctypes.c_void_p(self.off_dev.data_ptr()), w, d,
            self.code, mcs, ms, eps, ctypes.c_void_p(self.labels.data_ptr()), ctypes.c_void_p(self.ws.data_ptr()),
            need, ctypes.c_void_p(self.status.data_ptr()), flags | _native.FLAG_NO_GRAPH))
        return True

    def time_ms(self, flags, repeats):
        torch = self.torch
        if not self.run(flags):
            return None, None
        torch.cuda.synchronize()
        labels = self.labels.cpu().numpy().copy()
        samples = []
        for _ in range(repeats):
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            self.run(flags)
            stop.record()
            stop.synchronize()
            samples.append(start.elapsed_time(stop))
        if int(self.status.item()) != 0:
            raise RuntimeError("autotune: a tuning window reported non-finite values")
        return float(np.median(samples)), labels


def tune(d=16, precision="float32", min_cluster_size=10, min_samples=5, epsilon=0.05, windows=256,
         repeats=3, buckets=BUCKETS, seed=0, log=print):
    """Measures every bucket and returns the tuning table (a dict)."""
    import torch

    table = {"d": d, "precision": precision, "min_cluster_size": min_cluster_size, "min_samples": min_samples,
             "cluster_selection_epsilon": epsilon, "windows_per_bucket": windows, "repeats": repeats,
             "device": torch.cuda.get_device_name(0), "library": int(_native.lib().hdb_version()),
             "created": time.strftime("%Y-%m-%d %H:%M:%S"), "buckets": []}
    for top in BUCKETS:
        if top not in buckets:
            continue
        wins = bucket_windows(top, d, windows, seed + top)
        wins = [w for w in wins if len(w) >= min_samples]
        call = _Call(wins, precision, min_cluster_size, min_samples, epsilon)
        base_ms, base_labels = call.time_ms(0, repeats)
        best, best_ms = 0, base_ms
        measured = {0: base_ms}
        for name, options in DIMENSIONS:
            for opt in options:
                flags = (best & ~_mask_of(name)) | opt
                if flags in measured:
                    continue
                ms, labels = call.time_ms(flags, repeats)
                measured[flags] = ms
                if ms is None:
                    continue
                if not np.array_equal(labels, base_labels):
                    raise RuntimeError(f"autotune: flags {flags:#x} changed the labels of bucket {top}")
                if ms < best_ms:
                    best, best_ms = flags, ms
        entry = {"top": top, "chosen_flags": best, "chosen": describe(best), "chosen_ms": best_ms,
                 "auto_ms": base_ms, "measurements": [{"flags": f, **describe(f), "ms": m}
                                                      for f, m in sorted(measured.items())]}
        table["buckets"].append(entry)
        log(f"bucket <= {top}: auto {base_ms:.3f} ms, chosen {describe(best)} {best_ms:.3f} ms")
    return table


def _mask_of(name):
    return {"vertices_per_thread": 0xF00, "gram_row_sets": 0x3000, "repack": 0xC000,
            "arena": _native.FLAG_GLOBAL_ARENA | _native.FLAG_NO_POINT_LDS}[name]


def flags_of(table):
    """Per-bucket flags (7 entries; 0 = automatic) of a table."""
    flags = [0] * len(BUCKETS)
    for entry in table["buckets"]:
        flags[BUCKETS.index(int(entry["top"]))] = int(entry["chosen_flags"]) & TUNE_MASK
    return flags


def apply(table):
    """Installs a table for the ragged calls of this process (None clears)."""
    lib = _native.lib()
    if table is None:
        _native.check(lib.hdb_ragged_set_tuning(None, 0))
        return
    arr = (ctypes.c_int32 * len(BUCKETS))(*flags_of(table))
    _native.check(lib.hdb_ragged_set_tuning(arr, len(BUCKETS)))


def load(path):
    with open(path) as f:
        table = json.load(f)
    for key in ("d", "precision", "min_samples", "buckets"):
        if key not in table:
            raise ValueError(f"tuning table {path} lacks {key}")
    flags_of(table)  # validates the bucket entries
    return table


def save(table, path):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--d", type=int, default=16)
    ap.add_argument("--precision", default="float32", choices=("float32", "float64"))
    ap.add_argument("--min-cluster-size", type=int, default=10)
    ap.add_argument("--min-samples", type=int, default=5)
    ap.add_argument("--epsilon", type=float, default=0.05)
    ap.add_argument("--windows", type=int, default=256, help="windows per bucket")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    table = tune(args.d, args.precision, args.min_cluster_size, args.min_samples, args.epsilon, args.windows,
                 args.repeats, log=lambda *a: print(*a, flush=True))
    out = args.out or table_path(args.d, args.precision, args.min_samples)
    save(table, out)
    print(f"tuning table written to {out}", flush=True)


if __name__ == "__main__":
    main()
