"""Seeded synthetic window generators for the benchmark and parity workloads.

Each generator returns float64 host arrays (quantise with ``astype`` for the
float32 path).  NumPy's PCG64 ``default_rng`` streams are deterministic for a
given NumPy version, so the same seed reproduces the same windows on every
machine of this image.

Workloads (all Euclidean, labels-only HDBSCAN):
  single_256x8     N=256, D=8: 5 isotropic blobs (std 0.8, centres U[-10,10]^8)
                   + 10 % uniform noise; mcs=5, ms=5, eps=0
  grid_1024x8      B windows N=1024, D=8: 4-12 blobs (std U[0.5,1.5]) + 10 % noise,
                   coordinates rounded to a 1/8 grid; mcs=2, ms=2, eps=0.1
  aniso_4096x32    B windows N=4096, D=32 float64: 8-20 anisotropic rotated
                   components (axis scales U[0.2,2.0]) + 5 % noise; mcs=20, ms=10
  ragged_d16       windows with N log-uniform in [64, 4096], D=16: 3-15 blobs
                   + 10 % noise; mcs=10, ms=5, eps=0.05
  embed_2048x128   N=2048, D=128 L2-normalised rank-8 mixture (16 components)
                   + isotropic noise 0.05, 3 % exact duplicates; mcs=15, ms=8
  mixture_1024x16  B windows N=1024, D=16: 3-15 blobs + 10 % noise; mcs=5,
                   ms=5, eps=0 (the throughput workload)
"""

from __future__ import annotations

import numpy as np

WORKLOADS = {
    "single_256x8": dict(n=256, d=8, batch=1, precision="float32",
                         min_cluster_size=5, min_samples=5, cluster_selection_epsilon=0.0),
    "grid_1024x8": dict(n=1024, d=8, batch=64, precision="float32",
                        min_cluster_size=2, min_samples=2, cluster_selection_epsilon=0.1),
    "aniso_4096x32": dict(n=4096, d=32, batch=32, precision="float64",
                          min_cluster_size=20, min_samples=10, cluster_selection_epsilon=0.0),
    "ragged_d16": dict(n=None, d=16, batch=1000, precision="float32",
                       min_cluster_size=10, min_samples=5, cluster_selection_epsilon=0.05),
    "embed_2048x128": dict(n=2048, d=128, batch=128, precision="float32",
                           min_cluster_size=15, min_samples=8, cluster_selection_epsilon=0.0),
    "mixture_1024x16": dict(n=1024, d=16, batch=256, precision="float32",
                            min_cluster_size=5, min_samples=5, cluster_selection_epsilon=0.0),
}


def _blobs(rng, n, d, k, stds, noise_fraction, centre_range=10.0):
    noise = int(round(n * noise_fraction))
    inside = n - noise
    centres = rng.uniform(-centre_range, centre_range, (k, d))
    owner = rng.integers(0, k, inside)
    pts = centres[owner] + rng.normal(size=(inside, d)) * np.asarray(stds)[owner, None]
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    pts = np.concatenate([pts, rng.uniform(lo, hi, (noise, d))])
    return pts[rng.permutation(n)]


def single_256x8(seed=0, n=256, d=8):
    rng = np.random.default_rng(seed)
    return _blobs(rng, n, d, 5, np.full(5, 0.8), 0.10)


def grid_1024x8(seed=0, batch=64, n=1024, d=8):
    rng = np.random.default_rng(seed)
    out = np.empty((batch, n, d))
    for b in range(batch):
        k = int(rng.integers(4, 13))
        out[b] = _blobs(rng, n, d, k, rng.uniform(0.5, 1.5, k), 0.10)
    return np.round(out * 8) / 8


def aniso_4096x32(seed=0, batch=32, n=4096, d=32):
    rng = np.random.default_rng(seed)
    out = np.empty((batch, n, d))
    for b in range(batch):
        k = int(rng.integers(8, 21))
        noise = int(round(n * 0.05))
        inside = n - noise
        owner = rng.integers(0, k, inside)
        centres = rng.uniform(-10, 10, (k, d))
        pts = np.empty((inside, d))
        for c in range(k):
            idx = np.flatnonzero(owner == c)
            q, _ = np.linalg.qr(rng.normal(size=(d, d)))
            scales = rng.uniform(0.2, 2.0, d)
            pts[idx] = centres[c] + (rng.normal(size=(len(idx), d)) * scales) @ q.T
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        pts = np.concatenate([pts, rng.uniform(lo, hi, (noise, d))])
        out[b] = pts[rng.permutation(n)]
    return out


def ragged_sizes(seed=0, count=1000, low=64, high=4096):
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(low), np.log(high), count))).astype(np.int64)


def ragged_d16(seed=0, count=1000, d=16):
    """List of [N_i, D] windows."""
    sizes = ragged_sizes(seed, count)
    rng = np.random.default_rng(seed + 1)
    windows = []
    for n in sizes:
        k = int(rng.integers(3, 16))
        windows.append(_blobs(rng, int(n), d, k, rng.uniform(0.5, 1.5, k), 0.10))
    return windows


def embed_2048x128(seed=0, batch=128, n=2048, d=128, components=16, rank=8):
    rng = np.random.default_rng(seed)
    out = np.empty((batch, n, d))
    for b in range(batch):
        means = rng.normal(size=(components, d))
        bases = rng.normal(size=(components, d, rank)) / np.sqrt(rank)
        owner = rng.integers(0, components, n)
        latent = rng.normal(size=(n, rank))
        pts = means[owner] + np.einsum("ndr,nr->nd", bases[owner], latent)
        pts += rng.normal(0, 0.05, (n, d))
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
        dup = int(round(n * 0.03))
        src = rng.integers(0, n, dup)
        dst = rng.choice(n, dup, replace=False)
        pts[dst] = pts[src]
        out[b] = pts
    return out


def mixture_1024x16(seed=0, batch=256, n=1024, d=16):
    rng = np.random.default_rng(seed)
    out = np.empty((batch, n, d))
    for b in range(batch):
        k = int(rng.integers(3, 16))
        out[b] = _blobs(rng, n, d, k, rng.uniform(0.5, 1.5, k), 0.10)
    return out
