"""GPU parity: libhdbscan_rocm.so against the CPU oracle (bitwise) and the
golden scikit-learn fixtures.  Every call goes through the C ABI."""

import numpy as np
import pytest

import gpu_helpers as G
from golden_io import records, same_partition, workload_windows
from oracle import oracle
from hdbscan_amd._native import FLAG_GRAM_CHUNKED, FLAG_GRAM_GRAPH, FLAG_ONE_STEP, FLAG_SERIAL_LINKAGE, FLAG_SPLIT_TREE

pytestmark = pytest.mark.gpu

PRECISIONS = ["float32", "float64"]
BITS = {"float32": np.uint32, "float64": np.uint64}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("size", [2, 33, 129, 1024])
@pytest.mark.parametrize("dims", [1, 8, 31])
def test_distances_bitwise(precision, size, dims):
    rng = np.random.default_rng(size + dims)
    x = rng.normal(size=(3, size, dims)).astype(precision)
    x[0, 0] = 0
    x[0, -1] = -0.0
    x[1, -1] = x[1, 0]
    x[2] *= 1e9
    if size > 4:
        x[1, 3] = x[1, 4] + np.asarray(1e-30, precision)
    got, status = G.distances(x, precision)
    assert status == 0
    for b in range(3):
        exp = oracle.distances(x[b], precision)
        np.testing.assert_array_equal(got[b].view(BITS[precision]), exp.view(BITS[precision]))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("dims", [3, 16, 40, 100])
@pytest.mark.parametrize("ms", [3, 5, 8, 16, 17, 64, 65, 100])
@pytest.mark.parametrize("flags", [0, 64])
def test_core_distances_bitwise(precision, dims, ms, flags):
    """flags 0: the MFMA Gram filter where it applies (float32, ms <= 16);
    flags 64: every distance exact."""
    rng = np.random.default_rng(ms * 7 + dims)
    n = 300
    x = rng.integers(-6, 7, (2, n, dims)).astype(precision) / 4
    x[1] = rng.normal(size=(n, dims))
    got, status = G.core(x, ms, precision, flags)
    assert status == 0
    for b in range(2):
        exp = oracle.core(oracle.distances(x[b], precision), ms, precision)
        np.testing.assert_array_equal(got[b].view(BITS[precision]), exp.view(BITS[precision]))


def gram_core_windows():
    """float32 windows for the Gram-filtered core distances: every shape
    class of the launch (one wave to sixteen, ragged last chunk and strip),
    wide dimensions, unit-norm embeddings with exact duplicates, and the
    stress windows of the distance prefilter."""
    rng = np.random.default_rng(4242)
    out = {f"stress-{k}": v for k, v in prefilter_windows("float32").items()}
    for n, d in [(3, 2), (5, 4), (17, 8), (77, 16), (1000, 16), (1024, 16), (600, 128), (333, 200), (130, 256),
                 (1500, 16), (2048, 8), (1300, 32)]:
        out[f"normal-{n}x{d}"] = rng.normal(size=(n, d))
    e = rng.normal(size=(700, 128))
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    e[rng.choice(700, 21, replace=False)] = e[rng.integers(0, 700, 21)]
    out["embed-dup"] = e
    out["tiny-scale"] = rng.normal(size=(200, 8)) * 1e-19
    out["subnormal"] = rng.integers(-4, 5, (100, 3)) * 2.0 ** -140
    return {k: v.astype(np.float32) for k, v in out.items()}


_GRAM_CORE = {}


def gram_core_cached():
    if not _GRAM_CORE:
        _GRAM_CORE.update(gram_core_windows())
    return _GRAM_CORE


@pytest.mark.parametrize("ms", [3, 4, 8, 9, 16])
@pytest.mark.parametrize("rsets,chunked", [(0, 0)] + [(r, FLAG_GRAM_CHUNKED) for r in range(4)])
def test_core_gram_filter_bitwise(ms, rsets, chunked):
    """rsets: flag bits 12..13, 16-row sets per wave of the chunked Gram
    kernel (0: by window size, four for 1025..2048 points; 1, 2, 3: one, two,
    four).  chunked 0: windows of up to 1024 points and 16 dimensions take the
    per-window kernel (column operands in LDS, norms in the MFMA's K slots)."""
    for name, x in gram_core_cached().items():
        if ms > len(x):
            continue
        got, status = G.core(x, ms, "float32", (rsets << 12) | chunked)
        assert status == 0, name
        exp = oracle.core(oracle.distances(x, "float32"), ms, "float32")
        np.testing.assert_array_equal(got[0].view(np.uint32), exp.view(np.uint32), err_msg=name)


@pytest.mark.parametrize("ms", [5, 16])
def test_core_window_rare_paths_bitwise(ms):
    """The per-window Gram kernel's rarely taken paths: scalar row loads (a
    16-dimensional input that is not 16-byte aligned), rows whose candidates
    overflow the window's pair array (clusters of 20 exact duplicates: more
    than 16 candidates per row on average, so later rows take the
    whole-row path), and a window of one repeated point (every row past 32
    candidates)."""
    rng = np.random.default_rng(77)
    cases = {
        "misaligned": (rng.normal(size=(1000, 16)), True),
        "dup20": (np.repeat(rng.normal(size=(50, 16)) * 4, 20, axis=0), False),
        "same": (np.ones((300, 16)), False),
        "dup20-d7": (np.repeat(rng.normal(size=(40, 7)), 20, axis=0), False),
    }
    for name, (x, mis) in cases.items():
        x = x.astype(np.float32)
        got, status = G.core(x, ms, "float32", 0, misalign=mis)
        assert status == 0, name
        exp = oracle.core(oracle.distances(x, "float32"), ms, "float32")
        np.testing.assert_array_equal(got[0].view(np.uint32), exp.view(np.uint32), err_msg=name)


def path_graph(weights, dtype):
    size = len(weights) + 1
    dist = np.full((size, size), size * 10, dtype=dtype)
    np.fill_diagonal(dist, 0)
    for i, w in enumerate(weights):
        dist[i, i + 1] = dist[i + 1, i] = w
    return dist


ARGSORT = records("argsort")


@pytest.mark.parametrize("index", range(len(ARGSORT)), ids=[r["name"] for r in ARGSORT])
def test_introsort_order_matches_numpy(index):
    rec = ARGSORT[index]
    precision = "float32" if rec["is32"] else "float64"
    w = rec["w64"].astype(precision)
    # Distances are never negative: signed fixtures are folded to |w| and
    # checked against the (separately pinned) oracle order instead.
    signed = bool((w < 0).any())
    w = np.abs(w)
    dist = path_graph(w, precision)
    _, status, ex = G.cluster_matrix(dist, None, 2, 0.0, precision)
    assert status == 0
    np.testing.assert_array_equal(ex[0]["edge_w"], w)
    expected = oracle.argsort(w) if signed else rec["order"]
    np.testing.assert_array_equal(ex[0]["order"], expected)


@pytest.mark.parametrize("size", [33, 129, 1024, 4096])
@pytest.mark.parametrize("depth", [-1, 0, 1, 3])
@pytest.mark.parametrize("serial", [0, 16])
def test_forced_sort_depth_matches_oracle(size, depth, serial):
    rng = np.random.default_rng(size)
    w = rng.integers(1, 4, size - 1).astype(np.float32)
    dist = path_graph(w, np.float32)
    _, status, ex = G.cluster_matrix(dist, None, 2, 0.0, "float32", depth=depth, flags=serial)
    assert status == 0
    np.testing.assert_array_equal(ex[0]["order"], oracle.argsort(w, depth=depth))


def prim_matrices(size, dtype):
    rng = np.random.default_rng(size)
    equal = np.ones((size, size), dtype)
    retained = np.full((size, size), 9, dtype)
    if size >= 4:
        for a, b, value in [(0, 1, 1), (0, 2, 2), (1, 3, 3), (2, 3, 3)]:
            retained[a, b] = retained[b, a] = value
    quantized = rng.integers(1, 7, size=(size, size)).astype(dtype)
    quantized = np.minimum(quantized, quantized.T)
    mats = np.stack([equal, retained, quantized])
    for m in mats:
        np.fill_diagonal(m, 0)
    core = np.zeros((3, size), dtype)
    core[2] = rng.integers(0, 4, size=size)
    return mats, core


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("size", [2, 255, 256, 257, 513, 1500, 4096])
@pytest.mark.parametrize("vpt", [0, 1, 2, 4, 8])
@pytest.mark.parametrize("one_step", [0, FLAG_ONE_STEP])
def test_matrix_pipeline_matches_oracle(precision, size, vpt, one_step):
    """Matrix-row Prim (speculative two-row rounds, or one vertex per step)
    on all-equal, retained-edge and quantized matrices: every tie."""
    mats, core = prim_matrices(size, precision)
    if vpt == 8 and one_step:
        with pytest.raises(RuntimeError, match="8 vertices per thread"):
            G.cluster_matrix(mats, core, 2, 0.0, precision, flags=(vpt << 8) | one_step)
        return
    labels, status, ex = G.cluster_matrix(mats, core, 2, 0.0, precision, flags=(vpt << 8) | one_step)
    assert status == 0
    for b in range(3):
        u, v, w = oracle.prim(mats[b], core[b], precision=precision)
        np.testing.assert_array_equal(ex[b]["edge_u"], u)
        np.testing.assert_array_equal(ex[b]["edge_v"], v)
        np.testing.assert_array_equal(ex[b]["edge_w"], w)
        t = oracle.tree(u, v, w, 2, 0.0, precision=precision)
        for k in ("order", "left", "right", "sizes", "parent", "selected", "point_parent"):
            np.testing.assert_array_equal(ex[b][k], t[k], err_msg=k)
        for k in ("distance", "birth", "stability"):
            np.testing.assert_array_equal(ex[b][k].view(BITS[precision]), t[k].view(BITS[precision]), err_msg=k)
        np.testing.assert_array_equal(labels[b], t["labels"])


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("size", [2, 3, 100, 1025, 4096])
@pytest.mark.parametrize("flags", [0, FLAG_SERIAL_LINKAGE, 1, 1 | FLAG_SERIAL_LINKAGE])
def test_linkage_forms_match_oracle(precision, size, flags):
    """The run-form linkage (default) and the one-wave form, LDS and device
    arena, on all-equal (one tie group of every edge), retained-edge and
    quantized matrices: left / right / sizes exactly the oracle's."""
    mats, core = prim_matrices(size, precision)
    labels, status, ex = G.cluster_matrix(mats, core, 2, 0.0, precision, flags=flags)
    assert status == 0
    for b in range(3):
        u, v, w = oracle.prim(mats[b], core[b], precision=precision)
        t = oracle.tree(u, v, w, 2, 0.0, precision=precision)
        for k in ("order", "left", "right", "sizes"):
            np.testing.assert_array_equal(ex[b][k], t[k], err_msg=k)
        np.testing.assert_array_equal(labels[b], t["labels"])


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("levels", [2, 5, 40])
def test_linkage_ties_from_points(precision, levels):
    """Mutual-reachability weights with many equal values (points on a coarse
    grid, duplicates): the weight groups of the run-form linkage."""
    rng = np.random.default_rng(levels)
    x = (rng.integers(0, levels, (3, 600, 3)) / levels).astype(precision)
    x[:, 500:] = x[:, :100]
    for ms in (2, 4):
        labels, status, ex = G.cluster_export(x, 3, ms, 0.0, precision)
        assert status == 0
        for b in range(3):
            d = oracle.distances(x[b], precision)
            c = oracle.core(d, ms, precision)
            u, v, w = oracle.prim(d, c, precision=precision)
            t = oracle.tree(u, v, w, 3, 0.0, precision=precision)
            for k in ("order", "left", "right", "sizes"):
                np.testing.assert_array_equal(ex[b][k], t[k], err_msg=k)
            np.testing.assert_array_equal(labels[b], t["labels"])


@pytest.mark.parametrize("size", [32, 128])
@pytest.mark.parametrize("within_pair", [0.0, 0.25])
@pytest.mark.parametrize("epsilon", [0.0, 0.1, 2.0])
@pytest.mark.parametrize("minimum", [2, 5, 2**31 - 1])
@pytest.mark.parametrize("arena", [0, 1])
def test_eom_fixtures_match_oracle(size, within_pair, epsilon, minimum, arena):
    groups = np.arange(size) // 2
    chain = (np.maximum(groups[:, None], groups[None, :]) + 1).astype(np.float32)
    chain[groups[:, None] == groups[None, :]] = within_pair
    equal = np.ones_like(chain)
    nested = np.full_like(chain, 10)
    for width, distance in ((8, 1), (4, 100 / 101), (2, 10 / 11)):
        membership = np.arange(size) // width
        nested[membership[:, None] == membership[None, :]] = distance
    mats = np.stack([chain, equal, nested])
    for m in mats:
        np.fill_diagonal(m, 0)
    labels, status, ex = G.cluster_matrix(mats, np.zeros((3, size), np.float32), minimum, epsilon, "float32",
                                          flags=arena)
    assert status == 0
    for b in range(3):
        u, v, w = oracle.prim(mats[b], np.zeros(size, np.float32))
        t = oracle.tree(u, v, w, minimum, epsilon)
        for k in ("parent", "selected", "point_parent"):
            np.testing.assert_array_equal(ex[b][k], t[k], err_msg=k)
        for k in ("birth", "stability"):
            np.testing.assert_array_equal(ex[b][k], t[k], err_msg=k)
        np.testing.assert_array_equal(labels[b], t["labels"])


CORPUS = records("corpus")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_corpus_labels_match_oracle(precision):
    mismatches = []
    for r in CORPUS:
        x = r["x"].astype(precision)
        got, status = G.cluster(x, r["mcs"], r["ms"], r["eps"], precision)
        exp, est = oracle.hdbscan(x, r["mcs"], r["ms"], r["eps"], precision=precision)
        assert status == est == 0, r["name"]
        if not np.array_equal(got, exp):
            mismatches.append(r["name"])
    assert not mismatches, mismatches[:20]


WORKLOADS = workload_windows()


@pytest.mark.parametrize("index", range(len(WORKLOADS)), ids=[w[0] for w in WORKLOADS])
def test_workload_windows(index):
    name, x, precision, mcs, ms, eps, sk_labels = WORKLOADS[index]
    got, status = G.cluster(x, mcs, ms, eps, precision)
    assert status == 0
    exp, _ = oracle.hdbscan(x, mcs, ms, eps, precision=precision)
    np.testing.assert_array_equal(got, exp)
    assert same_partition(got, sk_labels)


@pytest.mark.parametrize("flags", [0, 1, 2, 4, 16, 32, 2 | 1, 1 << 8, 2 << 8, 4 << 8, 32 | 4 << 8,
                                   2 | FLAG_ONE_STEP, 2 | 4 << 8, 2 | 8 << 8,
                                   FLAG_ONE_STEP, 1 << 14, 2 << 14, 3 << 14, FLAG_ONE_STEP | 2 << 14,
                                   1 << 8 | 2 << 14, 4 << 8 | 3 << 14, 2 | FLAG_GRAM_GRAPH,
                                   2 | FLAG_GRAM_GRAPH | 4 << 8, FLAG_SPLIT_TREE, FLAG_SPLIT_TREE | 1,
                                   FLAG_SPLIT_TREE | 16, FLAG_SPLIT_TREE | 2, FLAG_SERIAL_LINKAGE,
                                   FLAG_SERIAL_LINKAGE | 1, FLAG_SPLIT_TREE | FLAG_SERIAL_LINKAGE])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_path_variants_agree(flags, precision):
    rng = np.random.default_rng(11)
    x = (rng.integers(-20, 21, (4, 700, 6)) / 8).astype(precision)
    for mcs, ms, eps in ((2, 2, 0.1), (5, 5, 0.0), (7, 3, 0.05)):
        got, status = G.cluster(x, mcs, ms, eps, precision, flags=flags)
        assert status == 0
        for b in range(4):
            exp, _ = oracle.hdbscan(x[b], mcs, ms, eps, precision=precision)
            np.testing.assert_array_equal(got[b], exp)


def test_status_flags():
    x = np.zeros((3, 33, 1), np.float32)
    x[-1, -1, 0] = np.finfo(np.float32).max
    assert G.cluster(x, 2, 2, 0.0, "float32")[1] == 1
    x[0, 0, 0] = np.finfo(np.float32).max
    x[-1, -1, 0] = np.nan
    assert G.cluster(x, 2, 2, 0.0, "float32")[1] & 2
    big = np.finfo(np.float64).max / 4
    assert G.cluster(np.array([[-big], [-big], [big], [big]]), 2, 5 - 3, 0.0, "float64")[1] == 1


def oracle_window(x, mcs, ms, eps, precision):
    """Oracle stages of one window: Prim edges and the tree intermediates."""
    dist = oracle.distances(x, precision)
    core = oracle.core(dist, ms, precision) if ms > 2 else None
    u, v, w = oracle.prim(dist, core, unit_core=ms <= 2, precision=precision)
    return (u, v, w), oracle.tree(u, v, w, mcs, eps, precision=precision)


def prefilter_windows(precision):
    """Windows that stress the certified distance prefilter: separations tiny
    against the norms (cancellation), ties, duplicates, norms just inside and
    beyond the limit under which the prefilter is enabled."""
    rng = np.random.default_rng(77)
    base = rng.normal(size=(256, 6))
    grid = rng.integers(-3, 4, (256, 6)) / 4.0
    out = {
        "offset-1e3": base + 1e3,
        "offset-1e5": base * 1e-3 + 1e5,
        "grid-ties": grid,
        "grid-offset": grid + 2 ** 20,
        "duplicates": np.repeat(base[:64], 4, axis=0),
        "huge-in": base * 2.0 ** 40,
        "huge-out": base * 2.0 ** 52,
    }
    if precision == "float64":
        out["huge-in"] = base * 2.0 ** 400
        out["huge-out"] = base * 2.0 ** 460
    return {k: v.astype(precision) for k, v in out.items()}


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("mcs,ms,eps", [(2, 2, 0.0), (5, 5, 0.0), (4, 3, 0.1)])
@pytest.mark.parametrize("vpt", [0, 1, 4])
@pytest.mark.parametrize("prefilter", [0, 32, FLAG_ONE_STEP])
def test_points_path_edges_bitwise(precision, mcs, ms, eps, vpt, prefilter):
    """prefilter 0: Prim from points; 32: with the certified Gram skip;
    FLAG_ONE_STEP: accepted (and without effect) on the points path."""
    for name, x in prefilter_windows(precision).items():
        labels, status, ex = G.cluster_export(x, mcs, ms, eps, precision, flags=(vpt << 8) | prefilter)
        assert status == 0, name
        (u, v, w), t = oracle_window(x, mcs, ms, eps, precision)
        np.testing.assert_array_equal(ex[0]["edge_u"], u, err_msg=name)
        np.testing.assert_array_equal(ex[0]["edge_v"], v, err_msg=name)
        np.testing.assert_array_equal(ex[0]["edge_w"].view(BITS[precision]), w.view(BITS[precision]), err_msg=name)
        for k in ("order", "left", "right", "sizes", "parent", "selected", "point_parent"):
            np.testing.assert_array_equal(ex[0][k], t[k], err_msg=f"{name}:{k}")
        for k in ("birth", "stability"):
            np.testing.assert_array_equal(ex[0][k].view(BITS[precision]), t[k].view(BITS[precision]),
                                          err_msg=f"{name}:{k}")
        np.testing.assert_array_equal(labels[0], t["labels"], err_msg=name)


def _edge_cases():
    """Every workload window through the default plan, and the windows of
    128 dimensions also with the MFMA Gram filter for the core distances
    (HDB_FLAG_MATRIX | HDB_FLAG_FILTER) and through the MFMA Gram graph
    stage (HDB_FLAG_GRAM_GRAPH)."""
    cases = [(i, 0) for i in range(len(WORKLOADS))]
    cases += [(i, f) for i, w in enumerate(WORKLOADS) if w[1].shape[1] == 128 for f in (2 | 128, FLAG_GRAM_GRAPH)]
    return cases


EDGE_CASES = _edge_cases()


@pytest.mark.parametrize("index,flags", EDGE_CASES, ids=[f"{WORKLOADS[i][0]}-f{f}" for i, f in EDGE_CASES])
def test_workload_edges_bitwise(index, flags):
    """Prim edges, tree intermediates and labels bitwise against the oracle's
    stages, at every workload size (the oracle takes a few seconds for the
    4096 x 32 float64 window)."""
    name, x, precision, mcs, ms, eps, _ = WORKLOADS[index]
    labels, status, ex = G.cluster_export(x, mcs, ms, eps, precision, flags=flags)
    assert status == 0
    (u, v, w), t = oracle_window(x, mcs, ms, eps, precision)
    np.testing.assert_array_equal(ex[0]["edge_u"], u)
    np.testing.assert_array_equal(ex[0]["edge_v"], v)
    np.testing.assert_array_equal(ex[0]["edge_w"].view(BITS[precision]), w.view(BITS[precision]))
    for k in ("order", "left", "right", "sizes", "parent", "selected", "point_parent"):
        np.testing.assert_array_equal(ex[0][k], t[k], err_msg=k)
    for k in ("birth", "stability"):
        np.testing.assert_array_equal(ex[0][k].view(BITS[precision]), t[k].view(BITS[precision]), err_msg=k)
    np.testing.assert_array_equal(labels[0], t["labels"])


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("ms", [3, 5, 9, 17, 40, 70])
@pytest.mark.parametrize("flags", [2 | 64, 2, 2 | 128, 2 | FLAG_GRAM_GRAPH, 2 | FLAG_GRAM_GRAPH | 1 << 8,
                                   2 | FLAG_GRAM_GRAPH | 2 << 8, 2 | FLAG_GRAM_GRAPH | 4 << 8])
def test_matrix_path_core_selection_edges(precision, ms, flags):
    """The exact matrix (flags 2, plain) with core distances selected from it
    or through the Gram filter (128), or, for float32, the MFMA Gram matrix
    with interval-key Prim and core distances from its bounds
    (HDB_FLAG_GRAM_GRAPH, any vertices per thread; the stress windows put
    separations far inside the bound, so many choices are resolved exactly;
    float64 ignores the flag).  Prim edge weights (which contain the core
    distances) bitwise."""
    for name, x in prefilter_windows(precision).items():
        labels, status, ex = G.cluster_export(x, 5, ms, 0.0, precision, flags=flags)
        assert status == 0, name
        (u, v, w), t = oracle_window(x, 5, ms, 0.0, precision)
        np.testing.assert_array_equal(ex[0]["edge_w"].view(BITS[precision]), w.view(BITS[precision]), err_msg=name)
        np.testing.assert_array_equal(ex[0]["edge_u"], u, err_msg=name)
        np.testing.assert_array_equal(labels[0], t["labels"], err_msg=name)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n", [2, 3, 5, 17, 64, 65, 200, 257, 1024, 1025, 2500, 4096])
def test_split_tree_stage_sizes(precision, n):
    """The split tree stage (sort kernel, sorted-edge transport, hierarchy
    kernel of max(64, (n-1)/4) threads) across window sizes: labels of the
    fused tree kernel and of the oracle."""
    rng = np.random.default_rng(n)
    x = (rng.integers(-12, 13, (3, n, 3)) / 4).astype(precision)
    ms = min(4, n)
    fused, st = G.cluster(x, 3, ms, 0.0, precision)
    assert st == 0
    split, st = G.cluster(x, 3, ms, 0.0, precision, flags=FLAG_SPLIT_TREE)
    assert st == 0
    np.testing.assert_array_equal(split, fused)
    for b in range(3):
        exp, _ = oracle.hdbscan(x[b], 3, ms, 0.0, precision=precision)
        np.testing.assert_array_equal(split[b], exp)
