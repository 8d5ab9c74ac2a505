"""Write the golden fixtures used by tests/test_oracle.py and the GPU parity tests.

Run in the build container (scikit-learn and NumPy importable):

    python tests/golden/make_golden.py

NumPy's SIMD argsort breaks ties differently from its scalar introsort, which
is the order the reference reproduces.  The script therefore re-runs itself in
a child process with NPY_DISABLE_CPU_FEATURES set before NumPy is imported.

Fixtures (inputs + expected outputs only):
  stages_f64.npz   float64 windows with scikit-learn's core distances, Prim MST,
                   single-linkage tree, condensed tree and labels
  corpus.npz       reference test corpora (float32-quantised inputs) with
                   scikit-learn labels for every parameter set the reference
                   tests use
  workloads.npz    windows of each benchmark workload with scikit-learn labels
  argsort.npz      tie-heavy arrays with np.argsort(kind="quicksort") orders
  heapsort.npz     the same arrays (sizes above 16) with np.argsort(kind="heapsort")
                   orders: NumPy's heapsort is the fallback of its introsort
                   once the depth budget is spent

    python tests/golden/make_golden.py heapsort   writes heapsort.npz only
"""

from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
from itertools import combinations_with_replacement

DISABLE = "AVX2 AVX512F AVX512CD AVX512_SKX AVX512_CLX AVX512_CNL AVX512_ICL AVX512_SPR"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def pack(records):
    """Concatenate a list of dicts of arrays (each flattened) with offsets."""
    import numpy as np

    out = {}
    keys = records[0].keys()
    for key in keys:
        arrays = [np.asarray(r[key]) for r in records]
        arrays = [a.ravel() if a.ndim > 1 else a for a in arrays]
        if arrays[0].ndim == 0:
            out[key] = np.stack(arrays)
            continue
        lengths = np.array([a.shape[0] for a in arrays], np.int64)
        out[key] = np.concatenate(arrays, axis=0) if lengths.sum() else arrays[0]
        out[key + "__offsets"] = np.concatenate([[0], np.cumsum(lengths)])
    return out


def sklearn_fit(x, **p):
    from sklearn.cluster import HDBSCAN

    return HDBSCAN(algorithm="auto", metric="euclidean", cluster_selection_method="eom",
                   allow_single_cluster=False, alpha=1.0, copy=True, **p).fit_predict(x)


def sklearn_stages(x, mcs, ms, eps):
    import numpy as np
    from sklearn.cluster._hdbscan._linkage import make_single_linkage, mst_from_data_matrix
    from sklearn.cluster._hdbscan._tree import _condense_tree
    from sklearn.metrics import DistanceMetric
    from sklearn.neighbors import NearestNeighbors

    x = np.ascontiguousarray(x, np.float64)
    nbrs = NearestNeighbors(n_neighbors=ms, algorithm="kd_tree", metric="euclidean").fit(x)
    core = np.ascontiguousarray(nbrs.kneighbors(x, ms, return_distance=True)[0][:, -1])
    metric = DistanceMetric.get_metric("euclidean")
    mst = mst_from_data_matrix(x, core, metric, 1.0)
    order = np.argsort(mst["distance"])
    slt = make_single_linkage(mst[order])
    condensed = _condense_tree(slt, mcs)
    labels = sklearn_fit(x, min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps)
    return {
        "x": x, "mcs": np.int64(mcs), "ms": np.int64(ms), "eps": np.float64(eps),
        "n": np.int64(x.shape[0]), "d": np.int64(x.shape[1]),
        "core": core,
        "mst_u": mst["current_node"].astype(np.int64), "mst_v": mst["next_node"].astype(np.int64),
        "mst_w": mst["distance"].astype(np.float64),
        "mst_order": order.astype(np.int64),
        "slt_left": slt["left_node"].astype(np.int64), "slt_right": slt["right_node"].astype(np.int64),
        "slt_value": slt["value"].astype(np.float64), "slt_size": slt["cluster_size"].astype(np.int64),
        "ct_parent": condensed["parent"].astype(np.int64), "ct_child": condensed["child"].astype(np.int64),
        "ct_value": condensed["value"].astype(np.float64), "ct_size": condensed["cluster_size"].astype(np.int64),
        "labels": labels.astype(np.int64),
    }


def seeded_sweep(seed):
    """The reference's seeded randomized sweep generator (tests/test_parity.py:175-195)."""
    import numpy as np

    rng = np.random.default_rng(seed + 1701)
    count = int(rng.integers(8, 81))
    dims = int(rng.integers(1, 9))
    if seed % 3 == 0:
        values = rng.integers(-8, 9, (count, dims)).astype(np.float32) / 4
    elif seed % 3 == 1:
        values = rng.normal(size=(count, dims)).astype(np.float32)
    else:
        centers = rng.normal(0, 4, (3, dims))
        values = (centers[np.arange(count) % 3] + rng.normal(0, 0.1, (count, dims))).astype(np.float32)
    mcs = int(rng.integers(2, min(10, count) + 1))
    ms = None if seed % 4 == 0 else int(rng.integers(1, min(8, count) + 1))
    return values, mcs, ms, 0.1 if seed % 2 else 0.0


def curated_corpus():
    """The reference's curated corpus (tests/test_parity.py:9-53)."""
    import numpy as np

    rng = np.random.default_rng(20261010)
    separated = np.concatenate([rng.normal(c, 0.08, (24, 2)) for c in ((-3, -3), (0, 3), (3, -3))])
    overlap = np.concatenate([rng.normal(c, 0.45, (30, 2)) for c in ((-0.5, 0), (0.5, 0), (0, 0.8))])
    varied = np.concatenate([rng.normal(c, s, (k, 2)) for c, s, k in
                             (((-4, 0), 0.025, 30), ((0, 4), 0.3, 35), ((4, 0), 0.8, 40))])
    grid = np.array([(x, y) for x in range(5) for y in range(5)])
    tiny = np.concatenate([rng.normal(c, 0.01, (k, 2)) for c, k in (((0, 0), 2), ((2, 2), 3), ((5, 0), 9))])
    return {
        "well-separated": separated.astype(np.float32),
        "overlap": overlap.astype(np.float32),
        "density-variation": varied.astype(np.float32),
        "random-quantized": rng.integers(-12, 13, (48, 3)).astype(np.float32) / 8,
        "duplicate-heavy": np.repeat(
            np.array([[0, 0], [0.125, 0], [4, 0], [4.125, 0], [20, 20]], np.float32), [7, 5, 6, 8, 1], axis=0),
        "all-identical": np.ones((24, 3), np.float32),
        "symmetric-grid-ties": grid.astype(np.float32),
        "symmetric-cross-ties": np.tile(np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1]], np.float32), (4, 1)),
        "tiny-clusters": tiny.astype(np.float32),
        "isolated-outliers": np.concatenate([separated, np.array([[20, 20], [-20, 20], [0, -30]], np.float32)]),
        "two-samples": np.array([[0, 0], [1, 1]], np.float32),
    }


def corpus_cases():
    """Every (window, parameters) pair of the reference's label parity tests."""
    import numpy as np

    cases = []
    corpus = curated_corpus()
    for name, values in corpus.items():
        for mcs, ms in ((2, 1), (2, 2), (5, None), (7, 3)):
            for eps in (0.0, 0.1):
                s = ms if ms is not None else mcs
                if s > len(values):
                    ms = len(values)
                cases.append((f"curated:{name}:{mcs}:{ms}:{eps}", values, mcs, ms, eps))
    grid = np.array([[0, 0], [0, 1], [1, 0], [1, 1]], np.float32)
    for size in (4, 5):
        for idx in combinations_with_replacement(range(4), size):
            for ms in (1, 2):
                for eps in (0.0, 0.1):
                    cases.append((f"grid:{''.join(map(str, idx))}:{ms}:{eps}", grid[list(idx)], 2, ms, eps))
    for idx in ((0, 0, 1, 2, 3), (0, 1, 1, 2, 2), (0, 0, 2, 3, 3)):
        for order in ((4, 3, 2, 1, 0), (2, 4, 0, 3, 1), (1, 0, 4, 2, 3)):
            cases.append((f"perm:{idx}:{order}", grid[list(idx)][list(order)], 2, 1, 0.0))
    for seed in range(36):
        values, mcs, ms, eps = seeded_sweep(seed)
        cases.append((f"sweep:{seed}", values, mcs, ms, eps))
    rng = np.random.default_rng(102408)
    centers = rng.normal(0, 5, (8, 8))
    real = (centers[np.arange(1024) % 8] + rng.normal(0, 0.2, (1024, 8))).astype(np.float32)
    cases.append(("real1024x8:5:None:0.0", real, 5, None, 0.0))
    cases.append(("real1024x8:12:6:0.1", real, 12, 6, 0.1))
    boundary = np.array([[0], [0], [0.125], [0.125], [4], [4], [4.125], [4.125]], np.float32)
    for eps in (0.0, 0.1, float(np.nextafter(np.float32(0.125), np.float32(0))), 0.125,
                float(np.nextafter(np.float32(0.125), np.float32(1)))):
        cases.append((f"boundary:{eps!r}", boundary, 2, 1, eps))
    for mcs, ms in ((2, 8), (8, 1), (9, 2)):
        cases.append((f"all-noise:{mcs}:{ms}", np.arange(16, dtype=np.float32).reshape(8, 2), mcs, ms, 0.0))
    rng = np.random.default_rng(31415)
    for i, w in enumerate([rng.normal(size=(32, 3)), rng.integers(-4, 5, (32, 3)) / 2, np.zeros((32, 3))]):
        for eps in (0.0, 0.1):
            cases.append((f"batch:{i}:{eps}", w.astype(np.float32), 3, 2, eps))
    return cases


def argsort_cases():
    import numpy as np

    records = []
    for size in (2, 5, 16, 17, 18, 19, 33, 129, 1024, 4096):
        rng = np.random.default_rng(size)
        m = size - 1
        patterns = {
            "equal": np.ones(m),
            "alternating": 1 + np.arange(m) % 2,
            "quantized": rng.integers(1, 5, m),
            "ascending": np.arange(1, size),
            "descending": np.arange(size - 1, 0, -1),
            "halves": np.repeat([2.0, 1.0], [m // 2, m - m // 2]),
            "random": rng.normal(size=m),
        }
        for name, w in patterns.items():
            for dt in (np.float32, np.float64):
                a = np.asarray(w, dt)
                records.append({"name": np.bytes_(f"{name}:{size}:{np.dtype(dt).name}"),
                                "w64": a.astype(np.float64), "is32": np.int64(dt == np.float32),
                                "order": np.argsort(a, kind="quicksort").astype(np.int64)})
    return records


def heapsort_cases():
    """NumPy's indirect heapsort on the argsort patterns (more than 16 keys:
    shorter spans never reach the heapsort fallback of the introsort)."""
    import numpy as np

    records = []
    for rec in argsort_cases():
        if len(rec["w64"]) <= 16:
            continue
        a = rec["w64"].astype(np.float32 if rec["is32"] else np.float64)
        records.append({"name": rec["name"], "w64": rec["w64"], "is32": rec["is32"],
                        "order": np.argsort(a, kind="heapsort").astype(np.int64)})
    return records


def workload_cases():
    import numpy as np

    sys.path.insert(0, ROOT)
    from hdbscan_amd import workloads as wl

    cases = []
    x = wl.single_256x8(seed=0)
    cases.append(("single_256x8:0", x.astype(np.float32), 5, 5, 0.0, "float32"))
    grid = wl.grid_1024x8(seed=0, batch=4)
    for b in range(4):
        cases.append((f"grid_1024x8:{b}", grid[b].astype(np.float32), 2, 2, 0.1, "float32"))
    mix = wl.mixture_1024x16(seed=0, batch=4)
    for b in range(4):
        cases.append((f"mixture_1024x16:{b}", mix[b].astype(np.float32), 5, 5, 0.0, "float32"))
    rag = wl.ragged_d16(seed=0, count=6)
    for b, w in enumerate(rag):
        cases.append((f"ragged_d16:{b}", w.astype(np.float32), 10, 5, 0.05, "float32"))
    an = wl.aniso_4096x32(seed=0, batch=1)
    cases.append(("aniso_4096x32:0", an[0], 20, 10, 0.0, "float64"))
    em = wl.embed_2048x128(seed=0, batch=1)
    cases.append(("embed_2048x128:0", em[0].astype(np.float32), 15, 8, 0.0, "float32"))
    return cases


def child():
    import numpy as np
    import sklearn

    np.savez_compressed(os.path.join(HERE, "heapsort.npz"), **pack(heapsort_cases()))
    if os.environ.get("HDB_GOLDEN_ONLY") == "heapsort":
        return

    info = {"sklearn": sklearn.__version__, "numpy": np.__version__,
            "NPY_DISABLE_CPU_FEATURES": os.environ.get("NPY_DISABLE_CPU_FEATURES")}
    # Stage fixtures: float64 windows (exactly representable float32 values
    # and genuinely float64 ones).
    stage_records = []
    for seed in range(36):
        values, mcs, ms, eps = seeded_sweep(seed)
        stage_records.append(sklearn_stages(values.astype(np.float64), mcs, ms if ms else mcs, eps))
    for seed in range(12):
        rng = np.random.default_rng(900 + seed)
        n = int(rng.integers(10, 200))
        d = int(rng.integers(1, 6))
        x = rng.normal(size=(n, d)) * (10.0 ** rng.integers(-3, 4))
        stage_records.append(sklearn_stages(x, int(rng.integers(2, 9)), int(rng.integers(1, 7)),
                                            float(rng.choice([0.0, 0.05, 0.5]))))
    for name, values in curated_corpus().items():
        if len(values) >= 5:
            stage_records.append(sklearn_stages(values.astype(np.float64), 5, 3, 0.0))
    np.savez_compressed(os.path.join(HERE, "stages_f64.npz"), **pack(stage_records))

    corpus_records = []
    for name, values, mcs, ms, eps in corpus_cases():
        values = np.asarray(values, np.float32)
        p = dict(min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps)
        labels = sklearn_fit(values, **p)
        corpus_records.append({"name": np.bytes_(name), "x": values.astype(np.float64),
                               "n": np.int64(values.shape[0]), "d": np.int64(values.shape[1]),
                               "mcs": np.int64(mcs), "ms": np.int64(mcs if ms is None else ms),
                               "eps": np.float64(eps), "labels": labels.astype(np.int64)})
    np.savez_compressed(os.path.join(HERE, "corpus.npz"), **pack(corpus_records))

    np.savez_compressed(os.path.join(HERE, "argsort.npz"), **pack(argsort_cases()))

    wl_records = []
    for name, values, mcs, ms, eps, precision in workload_cases():
        labels = sklearn_fit(values, min_cluster_size=mcs, min_samples=ms, cluster_selection_epsilon=eps)
        is32 = precision == "float32"
        wl_records.append({"name": np.bytes_(name),
                           "x32": np.asarray(values, np.float32) if is32 else np.zeros((0,), np.float32),
                           "x64": np.zeros((0,), np.float64) if is32 else np.asarray(values, np.float64),
                           "n": np.int64(values.shape[0]), "d": np.int64(values.shape[1]),
                           "is32": np.int64(precision == "float32"),
                           "mcs": np.int64(mcs), "ms": np.int64(ms), "eps": np.float64(eps),
                           "labels": labels.astype(np.int64),
                           "sha256": np.bytes_(hashlib.sha256(np.ascontiguousarray(values).tobytes()).hexdigest())})
        print("workload", name, values.shape, "clusters", labels.max() + 1, flush=True)
    np.savez_compressed(os.path.join(HERE, "workloads.npz"), **pack(wl_records))
    with open(os.path.join(HERE, "versions.json"), "w") as f:
        json.dump(info, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    if os.environ.get("HDB_GOLDEN_CHILD") == "1":
        child()
    else:
        env = dict(os.environ, NPY_DISABLE_CPU_FEATURES=DISABLE, HDB_GOLDEN_CHILD="1")
        if sys.argv[1:] == ["heapsort"]:
            env["HDB_GOLDEN_ONLY"] = "heapsort"
        sys.exit(subprocess.run([sys.executable, os.path.abspath(__file__)], env=env).returncode)
