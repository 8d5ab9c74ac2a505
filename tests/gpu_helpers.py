"""Calls into libhdbscan_rocm.so through its C ABI for the GPU parity tests."""

from __future__ import annotations

import ctypes

import numpy as np

from hdbscan_amd import _native

P = ctypes.c_void_p


def torch():
    import torch as t

    return t


def dev(a, dtype=None):
    """Host array -> device tensor through a freshly allocated pinned buffer
    (pageable copies of large arrays are not available on every host)."""
    t = torch()
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    pinned = t.from_numpy(arr).pin_memory()
    out = pinned.to("cuda", non_blocking=True)
    t.cuda.synchronize()
    return out


def host(x):
    t = torch()
    t.cuda.synchronize()
    pinned = t.empty(x.shape, dtype=x.dtype, pin_memory=True)
    pinned.copy_(x)
    t.cuda.synchronize()
    return pinned.numpy().copy()


def ptr(t):
    return None if t is None else P(t.data_ptr())


def stream():
    return P(torch().cuda.current_stream().cuda_stream)


def code(precision):
    return _native.F32 if precision == "float32" else _native.F64


def real(precision):
    return np.float32 if precision == "float32" else np.float64


def distances(x, precision):
    t = torch()
    x = np.asarray(x, real(precision))
    if x.ndim == 2:
        x = x[None]
    b, n, d = x.shape
    xd = dev(x)
    out = t.empty((b, n, n), dtype=xd.dtype, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    _native.check(_native.lib().hdb_distances(stream(), ptr(xd), b, n, d, code(precision), ptr(out), ptr(status)))
    return host(out), int(host(status)[0])


def bounded_distances(x, min_samples):
    """hdb_bounded_distances (float32): [B, N, N] exact distances or minus
    certified lower bounds, and the status."""
    t = torch()
    x = np.asarray(x, np.float32)
    if x.ndim == 2:
        x = x[None]
    b, n, d = x.shape
    lib = _native.lib()
    ws_bytes = lib.hdb_workspace_bytes(b, n, d, _native.F32, int(min_samples), _native.FLAG_MATRIX)
    _native.check(0 if ws_bytes else -1)
    xd = dev(x)
    ws = t.empty(ws_bytes, dtype=t.uint8, device="cuda")
    out = t.empty((b, n, n), dtype=t.float32, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    _native.check(lib.hdb_bounded_distances(stream(), ptr(xd), b, n, d, _native.F32, int(min_samples), ptr(out),
                                            ptr(ws), ws_bytes, ptr(status)))
    return host(out), int(host(status)[0])


def core(x, min_samples, precision, flags=0, misalign=False):
    """misalign: the points start one element past an allocation's start
    (a pointer that is not 16-byte aligned)."""
    t = torch()
    x = np.asarray(x, real(precision))
    if x.ndim == 2:
        x = x[None]
    b, n, d = x.shape
    xd = dev(x)
    xptr = ptr(xd)
    if misalign:
        flat = t.empty(x.size + 1, dtype=xd.dtype, device="cuda")
        flat[1:] = xd.reshape(-1)
        xptr = P(flat.data_ptr() + flat.element_size())
    out = t.empty((b, n), dtype=xd.dtype, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    ws_bytes = _native.lib().hdb_workspace_bytes(b, n, d, code(precision), int(min_samples),
                                                 int(flags) | 2 | _native.FLAG_NO_GRAM_EXACT
                                                 | (0 if flags & 64 else 128))
    assert ws_bytes > 0, _native.lib().hdb_last_error()
    ws = t.empty(ws_bytes, dtype=t.uint8, device="cuda")
    _native.check(_native.lib().hdb_core_distances(stream(), xptr, b, n, d, code(precision), int(min_samples),
                                                   ptr(out), ptr(ws), ws_bytes, ptr(status), int(flags)))
    return host(out), int(host(status)[0])


def cluster(x, mcs, ms, eps, precision, flags=0):
    """Labels [B, N] (numpy) and status from hdb_cluster."""
    t = torch()
    x = np.asarray(x, real(precision))
    squeeze = x.ndim == 2
    if squeeze:
        x = x[None]
    b, n, d = x.shape
    lib = _native.lib()
    ws_bytes = lib.hdb_workspace_bytes(b, n, d, code(precision), int(ms), int(flags))
    _native.check(0 if ws_bytes else -1)
    xd = dev(x)
    ws = t.empty(ws_bytes, dtype=t.uint8, device="cuda")
    labels = t.full((b, n), -7, dtype=t.int32, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    _native.check(lib.hdb_cluster(stream(), ptr(xd), b, n, d, code(precision), int(mcs), int(ms), float(eps),
                                  ptr(labels), ptr(ws), ws_bytes, ptr(status), int(flags)))
    out = host(labels)
    return (out[0] if squeeze else out), int(host(status)[0])


def cluster_export(x, mcs, ms, eps, precision, depth=None, flags=0):
    """hdb_cluster_export: labels, status and per-window intermediates."""
    t = torch()
    x = np.asarray(x, real(precision))
    if x.ndim == 2:
        x = x[None]
    b, n, d = x.shape
    lib = _native.lib()
    ws_bytes = lib.hdb_workspace_bytes(b, n, d, code(precision), int(ms), int(flags))
    _native.check(0 if ws_bytes else -1)
    xd = dev(x)
    ws = t.empty(ws_bytes, dtype=t.uint8, device="cuda")
    labels = t.full((b, n), -7, dtype=t.int32, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    ei = t.zeros((b, 10, n), dtype=t.int32, device="cuda")
    er = t.zeros((b, 4, n), dtype=xd.dtype, device="cuda")
    _native.check(lib.hdb_cluster_export(stream(), ptr(xd), b, n, d, code(precision), int(mcs), int(ms),
                                         float(eps), -100000 if depth is None else int(depth), ptr(labels),
                                         ptr(ws), ws_bytes, ptr(status), int(flags), ptr(ei), ptr(er)))
    return host(labels), int(host(status)[0]), unpack_exports(host(ei), host(er), n)


def unpack_exports(hi, hr, n):
    exports = []
    for w in range(hi.shape[0]):
        e = {k: hi[w, i] for i, k in enumerate(EXPORT_INTS)}
        e.update({k: hr[w, i] for i, k in enumerate(EXPORT_REALS)})
        nl = int(e["meta"][0])
        e["next_label"] = nl
        for k in ("edge_u", "edge_v", "order", "left", "right", "sizes", "edge_w", "distance"):
            e[k] = e[k][: n - 1]
        for k in ("parent", "selected", "birth", "stability"):
            e[k] = e[k][: nl - n]
        exports.append(e)
    return exports


EXPORT_INTS = ("edge_u", "edge_v", "order", "left", "right", "sizes", "parent", "selected", "point_parent", "meta")
EXPORT_REALS = ("edge_w", "distance", "birth", "stability")


def cluster_matrix(dist, core_dist, mcs, eps, precision, depth=None, flags=0, export=True):
    """Graph + tree stages from matrices; returns (labels, status, exports per window)."""
    t = torch()
    dist = np.asarray(dist, real(precision))
    if dist.ndim == 2:
        dist = dist[None]
    b, n, _ = dist.shape
    unit = core_dist is None
    lib = _native.lib()
    ws_bytes = lib.hdb_workspace_bytes(b, n, 1, code(precision), 1 if unit else 3,
                                       int(flags) | _native.FLAG_MATRIX | _native.FLAG_MATRIX_INPUT)
    _native.check(0 if ws_bytes else -1)
    dd = dev(dist)
    cd = None if unit else dev(np.asarray(core_dist, real(precision)).reshape(b, n))
    ws = t.empty(ws_bytes, dtype=t.uint8, device="cuda")
    labels = t.full((b, n), -7, dtype=t.int32, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    ei = t.zeros((b, 10, n), dtype=t.int32, device="cuda") if export else None
    er = t.zeros((b, 4, n), dtype=dd.dtype, device="cuda") if export else None
    _native.check(lib.hdb_cluster_matrix(stream(), ptr(dd), ptr(cd), b, n, code(precision), int(mcs), int(unit),
                                         float(eps), -100000 if depth is None else int(depth), ptr(labels),
                                         ptr(ws), ws_bytes, ptr(status), int(flags), ptr(ei), ptr(er)))
    out = host(labels)
    exports = unpack_exports(host(ei), host(er), n) if export else []
    return out, int(host(status)[0]), exports
