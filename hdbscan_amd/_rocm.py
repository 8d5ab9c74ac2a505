"""ROCm backend: the MI355X replacement of the reference's CUDA backend.

``cluster`` keeps the reference backend contract
(reference hdbscan_gpu/_cuda.py:91-128): it takes an already validated host
array or a native device array (here a ROCm ``torch.Tensor``), returns a newly
allocated device-resident ``int32`` label tensor ``[B, N]`` (``[N]`` for 2-D
input), enqueues all work on the caller's current stream and raises the
reference's ``ValueError`` messages.  It reads one status word back (one host
synchronisation) instead of the reference's two.  Labels are computed only by
libhdbscan_rocm.so; there is no CPU path.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native

_DTYPES = {"float32": _native.F32, "float64": _native.F64}


def _torch():
    import torch

    return torch


def _device_ready():
    torch = _torch()
    if not torch.cuda.is_available() or torch.cuda.device_count() < 1:
        raise RuntimeError("the ROCm backend requires a ROCm GPU")
    _native.lib()
    return torch


def _as_device(x, precision):
    """Converts to a contiguous device tensor of the working precision."""
    torch = _torch()
    dtype = torch.float32 if precision == "float32" else torch.float64
    if isinstance(x, torch.Tensor):
        if x.dtype == torch.bool or x.is_complex() or not (x.is_floating_point() or _is_int(x)):
            raise ValueError("X must contain real numeric values")
        device = x.device if x.is_cuda else torch.device("cuda", torch.cuda.current_device())
        return x.to(device=device, dtype=dtype).contiguous()
    host = np.ascontiguousarray(np.asarray(x), dtype=np.float32 if precision == "float32" else np.float64)
    # Stage through a freshly allocated pinned buffer: the copy is then a
    # plain DMA from page-locked memory on every host.
    # (pin_memory copies; a read-only array is copied first for from_numpy.)
    pinned = torch.from_numpy(host if host.flags.writeable else host.copy()).pin_memory()
    out = pinned.to(device=torch.device("cuda", torch.cuda.current_device()), non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return out


def _is_int(t):
    torch = _torch()
    return t.dtype in (torch.int8, torch.int16, torch.int32, torch.int64, torch.uint8)


def _stream_handle(device):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def raise_status(status):
    if status & _native.STATUS_NONFINITE:
        raise ValueError("X must contain only finite values in the requested precision")
    if status & _native.STATUS_OVERFLOW:
        raise ValueError("pairwise distances overflow the requested precision")


def cluster(x, min_cluster_size, min_samples, epsilon, precision, flags=0):
    """Labels of every window of ``x`` ([N, D] or [B, N, D]).

    Backend-level entry: callers (the public ``hdbscan`` function, the
    benchmark) are responsible for the admission checks of the public API.
    """
    torch = _device_ready()
    x = _as_device(x, precision)
    unbatched = x.dim() == 2
    if unbatched:
        x = x.unsqueeze(0)
    batch, n, d = (int(s) for s in x.shape)
    lib = _native.lib()
    dtype = _DTYPES[precision]
    ws_bytes = lib.hdb_workspace_bytes(batch, n, d, dtype, int(min_samples), int(flags))
    _native.check(0 if ws_bytes else -1)
    device = x.device
    with torch.cuda.device(device):
        workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
        status = torch.zeros(1, dtype=torch.int32, device=device)
        labels = torch.empty((batch, n), dtype=torch.int32, device=device)
        rc = lib.hdb_cluster(
            _stream_handle(device), ctypes.c_void_p(x.data_ptr()), batch, n, d, dtype,
            int(min_cluster_size), int(min_samples), float(epsilon),
            ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(workspace.data_ptr()), ws_bytes,
            ctypes.c_void_p(status.data_ptr()), int(flags))
        _native.check(rc)
        raise_status(int(status.item()))
    return labels[0] if unbatched else labels


def cluster_async(x, min_cluster_size, min_samples, epsilon, precision, flags=0, out=None,
                  workspace=None, status=None):
    """Enqueues one clustering of a contiguous device tensor without any sync.

    Returns (labels, status); the caller synchronises and checks ``status``
    with ``raise_status``.  Buffers may be passed in to reuse allocations.
    """
    torch = _torch()
    batch, n, d = (int(s) for s in x.shape)
    lib = _native.lib()
    dtype = _DTYPES[precision]
    ws_bytes = lib.hdb_workspace_bytes(batch, n, d, dtype, int(min_samples), int(flags))
    _native.check(0 if ws_bytes else -1)
    device = x.device
    if workspace is None or workspace.numel() < ws_bytes:
        workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    if status is None:
        status = torch.zeros(1, dtype=torch.int32, device=device)
    if out is None:
        out = torch.empty((batch, n), dtype=torch.int32, device=device)
    rc = lib.hdb_cluster(
        _stream_handle(device), ctypes.c_void_p(x.data_ptr()), batch, n, d, dtype,
        int(min_cluster_size), int(min_samples), float(epsilon),
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(workspace.data_ptr()), ws_bytes,
        ctypes.c_void_p(status.data_ptr()), int(flags))
    _native.check(rc)
    return out, status, workspace


def pack_windows(windows, precision):
    """Packs a list of [N_i, D] windows into one device tensor [sum N_i, D].

    Returns (packed, offsets) with ``offsets`` a host int64 array of length
    len(windows) + 1.
    """
    torch = _device_ready()
    if len(windows) == 0:
        raise ValueError("X requires nonempty batches/features and at least 2 samples")
    parts = [_as_device(w, precision) for w in windows]
    d = None
    for p in parts:
        if p.dim() != 2:
            raise ValueError("X must have rank 2 per window")
        if d is None:
            d = int(p.shape[1])
        elif int(p.shape[1]) != d:
            raise ValueError("every window must have the same number of features")
    offsets = np.zeros(len(parts) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([int(p.shape[0]) for p in parts])
    return torch.cat(parts, dim=0).contiguous(), offsets


def cluster_ragged_async(packed, offsets, min_cluster_size, min_samples, epsilon, precision, flags=0,
                         out=None, workspace=None, status=None, offsets_dev=None):
    """Enqueues one clustering of windows of uneven size (``packed`` [total, D]
    on the device, ``offsets`` the host window offsets) without any sync.
    Returns (labels [total], status, workspace, offsets_dev)."""
    torch = _torch()
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    windows = len(offsets) - 1
    total, d = (int(s) for s in packed.shape)
    lib = _native.lib()
    dtype = _DTYPES[precision]
    off_ptr = offsets.ctypes.data_as(ctypes.c_void_p)
    ws_bytes = lib.hdb_ragged_workspace_bytes(off_ptr, windows, d, dtype, int(min_samples), int(flags))
    _native.check(0 if ws_bytes else -1)
    device = packed.device
    if offsets_dev is None:
        offsets_dev = torch.from_numpy(offsets).to(device)
    if workspace is None or workspace.numel() < ws_bytes:
        workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    if status is None:
        status = torch.zeros(1, dtype=torch.int32, device=device)
    if out is None:
        out = torch.empty(total, dtype=torch.int32, device=device)
    rc = lib.hdb_cluster_ragged(
        _stream_handle(device), ctypes.c_void_p(packed.data_ptr()), off_ptr,
        ctypes.c_void_p(offsets_dev.data_ptr()), windows, d, dtype, int(min_cluster_size), int(min_samples),
        float(epsilon), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(workspace.data_ptr()), ws_bytes,
        ctypes.c_void_p(status.data_ptr()), int(flags))
    _native.check(rc)
    return out, status, workspace, offsets_dev


def cluster_ragged(windows, min_cluster_size, min_samples=None, cluster_selection_epsilon=0.0,
                   precision="float32", flags=0, tuning=None):
    """Labels of windows of uneven size in one call: a list of device int32
    label tensors, one per window ([N_i]).  Backend-level: the reference's
    public API is dense only (reference hdbscan_gpu/_api.py:71-78).
    ``tuning``: a table of ``hdbscan_amd.autotune`` (dict or JSON path) to
    install for this and later ragged calls."""
    torch = _device_ready()
    if min_samples is None:
        min_samples = min_cluster_size
    if tuning is not None:
        from . import autotune

        autotune.apply(autotune.load(tuning) if isinstance(tuning, str) else tuning)
    packed, offsets = pack_windows(windows, precision)
    labels, status, _, _ = cluster_ragged_async(packed, offsets, min_cluster_size, min_samples,
                                                cluster_selection_epsilon, precision, flags)
    raise_status(int(status.item()))
    return [labels[offsets[i]:offsets[i + 1]] for i in range(len(offsets) - 1)]
