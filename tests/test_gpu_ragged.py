"""GPU parity of ragged batches (windows of uneven size in one call) against
the CPU oracle, window by window; every call goes through the C ABI
(hdb_cluster_ragged)."""

import concurrent.futures as cf
import ctypes

import numpy as np
import pytest

import gpu_helpers as G
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")


def blobs(rng, n, d, precision):
    k = int(rng.integers(1, 6))
    centres = rng.uniform(-8, 8, (k, d))
    owner = rng.integers(0, k, n)
    x = centres[owner] + rng.normal(size=(n, d)) * rng.uniform(0.3, 1.5)
    if n > 8:
        x[: n // 10] = rng.uniform(-10, 10, (n // 10, d))
    return x.astype(precision)


def ragged(x_list, mcs, ms, eps, precision, flags=0):
    """Labels per window and the status word, through hdb_cluster_ragged."""
    torch = G.torch()
    from hdbscan_amd import _native

    offsets = np.zeros(len(x_list) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(w) for w in x_list])
    packed = G.dev(np.concatenate(x_list, axis=0))
    d = int(packed.shape[1])
    lib = _native.lib()
    code = G.code(precision)
    off_ptr = offsets.ctypes.data_as(ctypes.c_void_p)
    ws_bytes = lib.hdb_ragged_workspace_bytes(off_ptr, len(x_list), d, code, int(ms), int(flags))
    assert ws_bytes > 0, lib.hdb_last_error()
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    off_dev = torch.from_numpy(offsets).cuda()
    labels = torch.full((int(offsets[-1]),), -7, dtype=torch.int32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    _native.check(lib.hdb_cluster_ragged(G.stream(), G.ptr(packed), off_ptr, G.ptr(off_dev), len(x_list), d, code,
                                         int(mcs), int(ms), float(eps), G.ptr(labels), G.ptr(ws), ws_bytes,
                                         G.ptr(status), int(flags)))
    out = G.host(labels)
    return [out[offsets[i]:offsets[i + 1]] for i in range(len(x_list))], int(G.host(status)[0])


def expected(x_list, mcs, ms, eps, precision):
    def one(w):
        labels, status = oracle.hdbscan(w, mcs, ms, eps, precision=precision)
        assert status == 0
        return labels

    with cf.ThreadPoolExecutor(8) as pool:
        return list(pool.map(one, x_list))


SIZES = [2, 3, 9, 64, 65, 100, 128, 129, 200, 256, 257, 511, 600, 1024, 1025, 1500, 2048, 2049, 3000, 4096]


@pytest.mark.parametrize("precision,d", [("float32", 16), ("float32", 5), ("float64", 8)])
@pytest.mark.parametrize("mcs,ms,eps", [(10, 5, 0.05), (2, 2, 0.0), (4, 3, 0.0), (5, 16, 0.0), (3, 1, 0.2)])
def test_ragged_matches_oracle(precision, d, mcs, ms, eps):
    rng = np.random.default_rng(len(SIZES) * d + ms)
    sizes = [s for s in SIZES if s >= ms]
    order = rng.permutation(len(sizes))
    x_list = [blobs(rng, sizes[i], d, precision) for i in order]
    got, status = ragged(x_list, mcs, ms, eps, precision)
    assert status == 0
    for i, (g, e) in enumerate(zip(got, expected(x_list, mcs, ms, eps, precision))):
        np.testing.assert_array_equal(g, e, err_msg=f"window {i} n={len(x_list[i])}")


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_ragged_equal_sizes_match_dense(precision):
    rng = np.random.default_rng(5)
    x = np.stack([blobs(rng, 300, 8, precision) for _ in range(6)])
    dense, st = G.cluster(x, 5, 5, 0.0, precision)
    assert st == 0
    got, status = ragged(list(x), 5, 5, 0.0, precision)
    assert status == 0
    for b in range(6):
        np.testing.assert_array_equal(got[b], dense[b])


def test_ragged_workload_sample():
    from hdbscan_amd import workloads

    spec = workloads.WORKLOADS["ragged_d16"]
    x_list = [w.astype(np.float32) for w in workloads.ragged_d16(seed=3, count=40)]
    mcs, ms, eps = spec["min_cluster_size"], spec["min_samples"], spec["cluster_selection_epsilon"]
    got, status = ragged(x_list, mcs, ms, eps, "float32")
    assert status == 0
    for i, (g, e) in enumerate(zip(got, expected(x_list, mcs, ms, eps, "float32"))):
        np.testing.assert_array_equal(g, e, err_msg=f"window {i} n={len(x_list[i])}")


@pytest.mark.parametrize("flags", [64, 1, 32, 16, 1 << 20, (1 << 20) | 1, (1 << 20) | (1 << 19)])
def test_ragged_path_flags(flags):
    rng = np.random.default_rng(flags)
    x_list = [blobs(rng, n, 16, "float32") for n in (70, 300, 1200, 2500, 130)]
    got, status = ragged(x_list, 6, 4, 0.0, "float32", flags)
    assert status == 0
    for g, e in zip(got, expected(x_list, 6, 4, 0.0, "float32")):
        np.testing.assert_array_equal(g, e)


def test_ragged_nonfinite_and_overflow_status():
    rng = np.random.default_rng(9)
    x_list = [blobs(rng, n, 4, "float32") for n in (50, 80, 300)]
    bad = [w.copy() for w in x_list]
    bad[1][7, 2] = np.nan
    got, status = ragged(bad, 5, 5, 0.0, "float32")
    assert status & 2
    big = [w.copy() for w in x_list]
    big[2][0] = 3e38
    got, status = ragged(big, 5, 5, 0.0, "float32")
    assert status & 1 and not status & 2


@pytest.mark.parametrize("sizes,ms,d,message", [
    ([10, 1, 20], 2, 4, "at least 2"),
    ([10, 5000], 2, 4, "4096"),
    ([10, 4, 20], 5, 4, "min_samples"),
])
def test_ragged_rejects(sizes, ms, d, message):
    from hdbscan_amd import _native

    offsets = np.zeros(len(sizes) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    lib = _native.lib()
    got = lib.hdb_ragged_workspace_bytes(offsets.ctypes.data_as(ctypes.c_void_p), len(sizes), d, 0, ms, 0)
    assert got == 0
    assert message in lib.hdb_last_error().decode()


def test_cluster_ragged_wrapper():
    """The backend-level helper packs host or device windows and returns one
    device label tensor per window."""
    torch = G.torch()
    from hdbscan_amd import _rocm

    rng = np.random.default_rng(11)
    x_list = [blobs(rng, n, 6, "float32") for n in (40, 700, 90)]
    mixed = [x_list[0], torch.from_numpy(x_list[1]).cuda(), x_list[2].astype(np.float64)]
    got = _rocm.cluster_ragged(mixed, 5, 4, 0.0, "float32")
    assert [int(g.numel()) for g in got] == [40, 700, 90]
    assert all(g.is_cuda and g.dtype == torch.int32 for g in got)
    for g, e in zip(got, expected(x_list, 5, 4, 0.0, "float32")):
        np.testing.assert_array_equal(g.cpu().numpy(), e)


MATRIX_SIZES = [40, 300, 1100, 1500, 2048, 2049, 3000, 4096]


@pytest.mark.parametrize("precision,d", [("float32", 32), ("float32", 100), ("float64", 16)])
@pytest.mark.parametrize("mcs,ms,eps", [(10, 5, 0.05), (2, 2, 0.0), (5, 40, 0.0)])
def test_ragged_matrix_buckets(precision, d, mcs, ms, eps):
    """Windows whose coordinates do not fit the registers of the points path
    (float32 above 16 dimensions over 1024 points, float64 above 8) take the
    materialised matrix per window inside the ragged call."""
    rng = np.random.default_rng(d + ms)
    x_list = [blobs(rng, n, d, precision) for n in rng.permutation(MATRIX_SIZES)]
    got, status = ragged(x_list, mcs, ms, eps, precision)
    assert status == 0
    for i, (g, e) in enumerate(zip(got, expected(x_list, mcs, ms, eps, precision))):
        np.testing.assert_array_equal(g, e, err_msg=f"window {i} n={len(x_list[i])}")


class RaggedCall:
    """One ragged problem with fixed device buffers, so that repeated calls
    have the same signature and replay the captured graph."""

    def __init__(self, x_list, mcs, ms, eps, precision):
        torch = G.torch()
        from hdbscan_amd import _native

        self.lib = _native.lib()
        self.offsets = np.zeros(len(x_list) + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum([len(w) for w in x_list])
        self.packed = G.dev(np.concatenate(x_list, axis=0))
        self.args = (len(x_list), int(self.packed.shape[1]), G.code(precision), int(mcs), int(ms), float(eps))
        self.off_ptr = self.offsets.ctypes.data_as(ctypes.c_void_p)
        self.ws_bytes = self.lib.hdb_ragged_workspace_bytes(self.off_ptr, len(x_list), self.args[1], self.args[2],
                                                            int(ms), 0)
        assert self.ws_bytes > 0
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device="cuda")
        self.off_dev = torch.from_numpy(self.offsets).cuda()
        self.labels = torch.empty(int(self.offsets[-1]), dtype=torch.int32, device="cuda")
        self.status = torch.zeros(1, dtype=torch.int32, device="cuda")

    def run(self, flags=0):
        from hdbscan_amd import _native

        w, d, code, mcs, ms, eps = self.args
        self.labels.fill_(-7)
        self.status.zero_()
        _native.check(self.lib.hdb_cluster_ragged(G.stream(), G.ptr(self.packed), self.off_ptr, G.ptr(self.off_dev),
                                                  w, d, code, mcs, ms, eps, G.ptr(self.labels), G.ptr(self.ws),
                                                  self.ws_bytes, G.ptr(self.status), int(flags)))
        return G.host(self.labels), int(G.host(self.status)[0])


def test_ragged_graph_replays_match_direct_launches():
    """The first call captures a HIP graph, later calls with the same buffers
    replay it: every replay, with other device work and synchronisations in
    between, gives the labels of the directly launched call."""
    torch = G.torch()
    from hdbscan_amd import _native

    rng = np.random.default_rng(21)
    x_list = [blobs(rng, n, 16, "float32") for n in (64, 70, 300, 900, 1100, 2100, 3000, 129)]
    call = RaggedCall(x_list, 10, 5, 0.05, "float32")
    direct, st = call.run(_native.FLAG_NO_GRAPH)
    assert st == 0
    for e, g in zip(expected(x_list, 10, 5, 0.05, "float32"),
                    [direct[call.offsets[i]:call.offsets[i + 1]] for i in range(len(x_list))]):
        np.testing.assert_array_equal(g, e)
    scratch = torch.randn(1 << 20, device="cuda")
    for rep in range(5):
        got, st = call.run()
        assert st == 0
        np.testing.assert_array_equal(got, direct, err_msg=f"replay {rep}")
        scratch.mul_(1.5).add_(1.0)  # unrelated device work between replays
        torch.cuda.synchronize()
        if rep == 2:
            again, st = call.run(_native.FLAG_NO_GRAPH)
            np.testing.assert_array_equal(again, direct)


def test_ragged_graph_on_a_side_stream():
    torch = G.torch()
    rng = np.random.default_rng(22)
    x_list = [blobs(rng, n, 8, "float32") for n in (90, 500, 1300)]
    call = RaggedCall(x_list, 5, 5, 0.0, "float32")
    base, _ = call.run()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(3):
            got, st = call.run()
            assert st == 0
            np.testing.assert_array_equal(got, base)


def test_tuning_table_changes_no_label():
    """Any installed tuning table gives the labels of the automatic choice."""
    from hdbscan_amd import _native, autotune

    rng = np.random.default_rng(31)
    x_list = [blobs(rng, n, 16, "float32") for n in (40, 120, 250, 500, 1000, 2000, 3500)]
    base, st = ragged(x_list, 10, 5, 0.05, "float32")
    assert st == 0
    flags = [(4 << 8) | (1 << 14), 2 << 8, (1 << 8) | _native.FLAG_GLOBAL_ARENA, 3 << 14,
             (3 << 12) | _native.FLAG_NO_POINT_LDS, (2 << 8) | (1 << 12), (4 << 8) | (2 << 14)]
    try:
        autotune.apply({"d": 16, "precision": "float32", "min_samples": 5,
                        "buckets": [{"top": t, "chosen_flags": f} for t, f in zip(autotune.BUCKETS, flags)]})
        got, st = ragged(x_list, 10, 5, 0.05, "float32")
    finally:
        autotune.apply(None)
    assert st == 0
    for g, b in zip(got, base):
        np.testing.assert_array_equal(g, b)


def test_autotune_small_run():
    from hdbscan_amd import autotune

    table = autotune.tune(d=8, windows=6, repeats=1, buckets=(128, 1024), log=lambda *a: None)
    assert [b["top"] for b in table["buckets"]] == [128, 1024]
    for b in table["buckets"]:
        assert b["chosen_ms"] <= b["auto_ms"]
        assert len(b["measurements"]) >= 5
