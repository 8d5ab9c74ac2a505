"""Readers for the committed golden fixtures in tests/golden/."""

from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def records(name):
    """Unpack tests/golden/<name>.npz into a list of per-case dicts."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as data:
        arrays = {k: data[k] for k in data.files}
    keys = [k for k in arrays if not k.endswith("__offsets")]
    count = None
    for k in keys:
        off = arrays.get(k + "__offsets")
        c = len(off) - 1 if off is not None else len(arrays[k])
        count = c if count is None else count
        assert count == c, k
    out = []
    for i in range(count):
        rec = {}
        for k in keys:
            off = arrays.get(k + "__offsets")
            if off is None:
                v = arrays[k][i]
                rec[k] = v.decode() if isinstance(v, bytes) else v.item() if np.ndim(v) == 0 else v
            else:
                rec[k] = arrays[k][off[i]:off[i + 1]]
        for xk in ("x", "x32", "x64"):
            if xk in rec and "n" in rec and rec[xk].size:
                rec[xk] = rec[xk].reshape(int(rec["n"]), int(rec["d"]))
        out.append(rec)
    return out


def workload_windows():
    """(name, x, precision, mcs, ms, eps, sklearn_labels) for every workload fixture."""
    out = []
    for r in records("workloads"):
        is32 = bool(r["is32"])
        x = r["x32"] if is32 else r["x64"]
        out.append((r["name"], x, "float32" if is32 else "float64", int(r["mcs"]), int(r["ms"]),
                    float(r["eps"]), r["labels"].astype(np.int32)))
    return out


def same_partition(actual, expected):
    """Exact noise mask plus a cluster bijection (reference tests/conftest.py:116-134)."""
    actual = np.asarray(actual).ravel()
    expected = np.asarray(expected).ravel()
    if actual.shape != expected.shape or not np.array_equal(actual == -1, expected == -1):
        return False
    forward, reverse = {}, {}
    for a, e in zip(actual.tolist(), expected.tolist()):
        if e == -1:
            continue
        if a < 0 or forward.setdefault(a, e) != e or reverse.setdefault(e, a) != a:
            return False
    return True
