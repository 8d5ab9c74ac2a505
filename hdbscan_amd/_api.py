"""Public entry point: ``hdbscan(X, ...)`` with ``backend="rocm"``.

Mirrors the reference's public function (reference hdbscan_gpu/_api.py:11-102):
the same keyword arguments, validation order, limits (4096 samples per
window, 512 MiB for the two dense distance buffers of the reference design)
and ValueError texts, so the reference's API tests read the same.  Native
device input is a ROCm ``torch.Tensor``; host input is anything
``np.asarray`` accepts (a CPU tensor is treated as host data).  The MLX and
CUDA backends of the reference are not part of this package.
"""

from __future__ import annotations

import math
from numbers import Integral, Real

import numpy as np

MAX_SAMPLES = 4096
MAX_DISTANCE_BYTES = 512 * 1024 * 1024
BACKENDS = ("rocm",)


def _is_native(X):
    module = type(X).__module__.split(".")[0]
    if module != "torch":
        return False
    return bool(getattr(X, "is_cuda", False))


def hdbscan(
    X,
    *,
    min_cluster_size=5,
    min_samples=None,
    cluster_selection_epsilon=0.0,
    backend="rocm",
    precision="float32",
):
    """Cluster each window of ``X`` on the GPU and return its labels.

    ``X`` is one window ``[N, D]`` or a batch ``[B, N, D]`` of independent
    windows, either a ROCm ``torch.Tensor`` (used in place, never copied to
    the host) or host data (NumPy arrays, nested lists, CPU tensors), which is
    copied to the current GPU once.  The clustering is Euclidean HDBSCAN with
    excess-of-mass selection; the root cluster is never selected, and a
    point's own position counts towards ``min_samples``.  The result is a new
    ``int32`` tensor ``[B, N]`` (``[N]`` for one window) on the same GPU,
    ``-1`` marking noise, enqueued on the current stream.  The only
    host synchronisation is one status word, which turns a non-finite input
    or an overflowing distance into a ``ValueError``.  Work happens in
    ``precision`` (float32 or float64); in float32, pairs whose distances tie
    or nearly tie can order differently than scikit-learn's float64 run.
    Limits: 4096 points per window, and ``2 * B * N**2 * itemsize`` at most
    512 MiB per call (the admission rule of the reference API).
    """
    if backend not in BACKENDS:
        raise ValueError("backend must be 'rocm'")
    if precision not in ("float32", "float64"):
        raise ValueError("precision must be 'float32' or 'float64'")
    for name, value, lower in (
        ("min_cluster_size", min_cluster_size, 2),
        ("min_samples", min_samples, 1),
    ):
        if value is None and name == "min_samples":
            continue
        if isinstance(value, (bool, np.bool_)) or not isinstance(value, Integral):
            raise ValueError(f"{name} must be an integer")
        if value < lower:
            raise ValueError(f"{name} must be >= {lower}")
    if (
        isinstance(cluster_selection_epsilon, (bool, np.bool_))
        or not isinstance(cluster_selection_epsilon, Real)
        or not math.isfinite(cluster_selection_epsilon)
        or cluster_selection_epsilon < 0
    ):
        raise ValueError("cluster_selection_epsilon must be finite and >= 0")
    if min_samples is None:
        min_samples = min_cluster_size

    native = _is_native(X)
    if native:
        if X.dtype.is_complex or str(X.dtype) == "torch.bool" or not (
                X.dtype.is_floating_point or str(X.dtype) in (
                    "torch.int8", "torch.int16", "torch.int32", "torch.int64", "torch.uint8")):
            raise ValueError("X must contain real numeric values")
        shape = tuple(X.shape)
    else:
        module = type(X).__module__.split(".")[0]
        if module == "torch":
            X = X.detach().numpy()
        elif hasattr(X, "__cuda_array_interface__") or (
                hasattr(X, "__dlpack_device__") and not isinstance(X, np.ndarray)):
            # Device arrays of other frameworks are not read through the host
            # (reference hdbscan_gpu/_api.py:62-64).
            raise TypeError("use a native ROCm torch tensor or a NumPy host array")
        X = np.asarray(X)
        if X.dtype.kind not in "fiu":
            raise ValueError("X must contain real numeric values")
        if not np.isfinite(X).all():
            raise ValueError("X must contain only finite values")
        shape = X.shape

    if len(shape) not in (2, 3):
        raise ValueError("X must have rank 2 [N,D] or rank 3 [B,N,D]")
    n, dimensions = shape[-2:]
    batch = shape[0] if len(shape) == 3 else 1
    if batch < 1 or dimensions < 1 or n < 2:
        raise ValueError("X requires nonempty batches/features and at least 2 samples")
    if n > MAX_SAMPLES:
        raise ValueError(f"at most {MAX_SAMPLES} samples per window are supported")
    if min_samples > n:
        raise ValueError("min_samples must not exceed the number of samples")
    itemsize = 4 if precision == "float32" else 8
    if 2 * batch * n * n * itemsize > MAX_DISTANCE_BYTES:
        raise ValueError("dense distance workspace exceeds the 512 MiB limit")
    if min_cluster_size > np.iinfo(np.int32).max:
        raise ValueError("min_cluster_size must fit in int32")
    if cluster_selection_epsilon > np.finfo(precision).max:
        raise ValueError("cluster_selection_epsilon is not representable in the requested precision")

    if not native:
        X = np.ascontiguousarray(X)
    from ._rocm import cluster

    return cluster(
        X,
        int(min_cluster_size),
        int(min_samples),
        float(cluster_selection_epsilon),
        precision,
    )
