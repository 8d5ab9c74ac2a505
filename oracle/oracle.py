"""ctypes binding of the CPU oracle (oracle/hdb_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of ``bench.py`` -- always as the checker or as the
timed CPU comparator, never as part of the product path.

Every function takes and returns NumPy arrays for ONE window unless named
otherwise.  ``precision`` selects the Real type ("float32" / "float64").
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

_I = ctypes.c_int
_LL = ctypes.c_longlong
_P = ctypes.c_void_p

_SIGNATURES = {
    "orc_distances": (None, [_P, _I, _I, _P]),
    "orc_core": (None, [_P, _I, _I, _P]),
    "orc_prim": (None, [_P, _P, _I, _I, _P, _P, _P]),
    "orc_argsort": (None, [_P, _I, _I, _P]),
    "orc_linkage": (None, [_P, _P, _P, _P, _I, _P, _P, _P, _P]),
    "orc_tree": (_I, [_P, _P, _P, _I, _LL, None, _I, _P] + [_P] * 10),
    "orc_hdbscan": (_I, [_P, _I, _I, _I, _LL, _I, None, _P]),
}


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        handle = ctypes.CDLL(_LIB_PATH)
        for base, (restype, argtypes) in _SIGNATURES.items():
            for sfx, real in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
                fn = getattr(handle, f"{base}_{sfx}")
                fn.restype = restype
                fn.argtypes = [real if a is None else a for a in argtypes]
        _lib = handle
    return _lib


def _sfx(precision):
    if precision in ("float32", np.float32):
        return "f32", np.float32
    if precision in ("float64", np.float64):
        return "f64", np.float64
    raise ValueError(precision)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def distances(x, precision="float32"):
    sfx, real = _sfx(precision)
    x = np.ascontiguousarray(x, dtype=real)
    n, d = x.shape
    out = np.empty((n, n), dtype=real)
    getattr(lib(), f"orc_distances_{sfx}")(_ptr(x), n, d, _ptr(out))
    return out


def core(dist, min_samples, precision="float32"):
    sfx, real = _sfx(precision)
    dist = np.ascontiguousarray(dist, dtype=real)
    n = dist.shape[0]
    out = np.empty(n, dtype=real)
    getattr(lib(), f"orc_core_{sfx}")(_ptr(dist), n, int(min_samples), _ptr(out))
    return out


def prim(dist, core_dist, unit_core=False, precision="float32"):
    """Prim edges (u, v, w) in step order."""
    sfx, real = _sfx(precision)
    dist = np.ascontiguousarray(dist, dtype=real)
    n = dist.shape[0]
    core_dist = np.ascontiguousarray(
        np.zeros(n) if core_dist is None else core_dist, dtype=real)
    u = np.empty(max(n - 1, 1), np.int32)
    v = np.empty(max(n - 1, 1), np.int32)
    w = np.empty(max(n - 1, 1), real)
    getattr(lib(), f"orc_prim_{sfx}")(_ptr(dist), _ptr(core_dist), n, int(bool(unit_core)),
                                      _ptr(u), _ptr(v), _ptr(w))
    return u[: n - 1], v[: n - 1], w[: n - 1]


def argsort(values, depth=None, precision=None):
    values = np.asarray(values)
    sfx, real = _sfx(precision or values.dtype.type)
    values = np.ascontiguousarray(values, dtype=real)
    m = values.shape[0]
    order = np.empty(max(m, 1), np.int32)
    getattr(lib(), f"orc_argsort_{sfx}")(_ptr(values), m,
                                         -100000 if depth is None else int(depth), _ptr(order))
    return order[:m]


def tree(edge_u, edge_v, edge_w, min_cluster_size, epsilon=0.0, depth=None, precision="float32"):
    """Sort, link, condense, select and label one window's Prim edges.

    Returns a dict with every intermediate array (cluster-indexed arrays are
    indexed by cluster_id - n; only the first next_label - n are meaningful).
    """
    sfx, real = _sfx(precision)
    n = len(edge_u) + 1
    eu = np.ascontiguousarray(edge_u, np.int32)
    ev = np.ascontiguousarray(edge_v, np.int32)
    ew = np.ascontiguousarray(edge_w, real)
    m = max(n - 1, 1)
    out = {
        "labels": np.empty(n, np.int32),
        "order": np.empty(m, np.int32),
        "left": np.empty(m, np.int32),
        "right": np.empty(m, np.int32),
        "sizes": np.empty(m, np.int32),
        "distance": np.empty(m, real),
        "parent": np.empty(n, np.int32),
        "birth": np.empty(n, real),
        "stability": np.empty(n, real),
        "selected": np.empty(n, np.uint8),
        "point_parent": np.empty(n, np.int32),
    }
    next_label = getattr(lib(), f"orc_tree_{sfx}")(
        _ptr(eu), _ptr(ev), _ptr(ew), n, int(min_cluster_size), float(epsilon),
        -100000 if depth is None else int(depth), _ptr(out["labels"]),
        *(_ptr(out[k]) for k in ("order", "left", "right", "sizes", "distance", "parent",
                                  "birth", "stability", "selected", "point_parent")))
    for k in ("order", "left", "right", "sizes", "distance"):
        out[k] = out[k][: n - 1]
    clusters = next_label - n
    for k in ("parent", "birth", "stability", "selected"):
        out[k] = out[k][:clusters]
    out["next_label"] = next_label
    return out


def hdbscan(x, min_cluster_size=5, min_samples=None, cluster_selection_epsilon=0.0,
            precision="float32"):
    """Full pipeline on [N, D] or [B, N, D]; returns (labels, status).

    status bit 1 = non-finite input, bit 0 = non-finite distance (labels are
    then undefined).
    """
    sfx, real = _sfx(precision)
    x = np.asarray(x)
    unbatched = x.ndim == 2
    x = np.ascontiguousarray(x[None] if unbatched else x, dtype=real)
    b, n, d = x.shape
    ms = min_cluster_size if min_samples is None else min_samples
    labels = np.full((b, n), -1, np.int32)
    status = getattr(lib(), f"orc_hdbscan_{sfx}")(
        _ptr(x), b, n, d, int(min_cluster_size), int(ms),
        float(real(cluster_selection_epsilon)), _ptr(labels))
    return (labels[0] if unbatched else labels), status
