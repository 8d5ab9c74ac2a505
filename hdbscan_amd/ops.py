"""``torch.ops.hdbscan_amd.cluster``: the ROCm backend as a PyTorch custom op.

Importing this module registers the op.  It takes a device tensor
``[N, D]`` or ``[B, N, D]`` (any real dtype, converted to the working
precision like the backend does) and returns the int32 labels ``[N]`` or
``[B, N]`` on the same device, enqueued on the current stream.  The fake
(meta) implementation gives the output's shape, dtype and device, so graphs
containing the op trace under ``torch.compile`` with ``fullgraph=True``; at
run time the op is opaque and raises the backend's ``ValueError``s
(non-finite input, overflow).  The reference has no such op: its public
function is plain Python (reference hdbscan_gpu/_api.py:30-78); argument
meanings follow that function, with ``min_samples`` already resolved.
"""

from __future__ import annotations

import math

import torch

from . import _rocm
from ._api import MAX_SAMPLES

PRECISIONS = ("float32", "float64")


@torch.library.custom_op("hdbscan_amd::cluster", mutates_args=(), device_types="cuda")
def cluster(x: torch.Tensor, min_cluster_size: int, min_samples: int, epsilon: float,
            precision: str) -> torch.Tensor:
    """Labels of every window of ``x`` (see the module docstring)."""
    _check(x, min_cluster_size, min_samples, epsilon, precision)
    labels = _rocm.cluster(x, min_cluster_size, min_samples, epsilon, precision)
    # A fresh tensor, never a view of the input.
    return labels.clone() if labels._base is not None else labels


@cluster.register_fake
def _cluster_fake(x, min_cluster_size, min_samples, epsilon, precision):
    _check(x, min_cluster_size, min_samples, epsilon, precision)
    return x.new_empty(x.shape[:-1], dtype=torch.int32)


def _check(x, min_cluster_size, min_samples, epsilon, precision):
    """The public function's checks that need no data (same messages)."""
    if precision not in PRECISIONS:
        raise ValueError("precision must be 'float32' or 'float64'")
    if min_cluster_size < 2:
        raise ValueError("min_cluster_size must be >= 2")
    if min_samples < 1:
        raise ValueError("min_samples must be >= 1")
    if not math.isfinite(epsilon) or epsilon < 0:
        raise ValueError("cluster_selection_epsilon must be finite and >= 0")
    if x.dtype.is_complex or x.dtype == torch.bool:
        raise ValueError("X must contain real numeric values")
    if x.dim() not in (2, 3):
        raise ValueError("X must have rank 2 [N,D] or rank 3 [B,N,D]")
    n, dimensions = x.shape[-2:]
    if (x.dim() == 3 and x.shape[0] < 1) or dimensions < 1 or n < 2:
        raise ValueError("X requires nonempty batches/features and at least 2 samples")
    if n > MAX_SAMPLES:
        raise ValueError(f"at most {MAX_SAMPLES} samples per window are supported")
    if min_samples > n:
        raise ValueError("min_samples must not exceed the number of samples")
