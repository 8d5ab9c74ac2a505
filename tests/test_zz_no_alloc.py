"""hdb_cluster_noalloc (the entry point that allocates nothing) against
hdb_cluster with a caller workspace: the same labels and status, on two
calls in a row, the second with a smaller input, so the second call runs on
scratch memory that still holds the first call's contents."""

import ctypes

import numpy as np
import pytest

import gpu_helpers as G
from hdbscan_amd import _native, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    t = pytest.importorskip("torch")
    if not t.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return t


def noalloc(x, mcs, ms, eps, precision, flags=0):
    t = G.torch()
    b, n, d = x.shape
    xd = G.dev(x)
    labels = t.full((b, n), -7, dtype=t.int32, device="cuda")
    status = t.zeros(1, dtype=t.int32, device="cuda")
    print(f"hdb_cluster_noalloc B={b} N={n} D={d} {precision} mcs={mcs} ms={ms} eps={eps} flags={flags}",
          flush=True)
    _native.check(_native.lib().hdb_cluster_noalloc(
        G.stream(), G.ptr(xd), b, n, d, G.code(precision), int(mcs), int(ms), float(eps), G.ptr(labels),
        G.ptr(status), int(flags)))
    return G.host(labels), int(G.host(status)[0])


CASES = [
    # (first call, second smaller call): inputs, parameters, precision
    ((workloads.mixture_1024x16(seed=3, batch=24), 5, 5, 0.0, "float32"),
     (workloads.grid_1024x8(seed=4, batch=2)[:, :700], 2, 2, 0.1, "float32")),
    ((workloads.embed_2048x128(seed=5, batch=2), 15, 8, 0.0, "float32"),
     (workloads.single_256x8(seed=6)[None], 5, 5, 0.0, "float32")),
    ((np.random.default_rng(7).normal(size=(4, 1500, 12)), 10, 7, 0.0, "float64"),
     (np.random.default_rng(8).normal(size=(1, 99, 3)), 3, 3, 0.05, "float64")),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_matches_workspace_entry(torch, case):
    for x, mcs, ms, eps, precision in CASES[case]:
        x = np.ascontiguousarray(x, G.real(precision))
        want, want_status = G.cluster(x, mcs, ms, eps, precision)
        got, status = noalloc(x, mcs, ms, eps, precision)
        assert status == want_status == 0
        np.testing.assert_array_equal(got, want)


def test_noalloc_status_and_limit(torch):
    x = np.zeros((2, 40, 3), np.float32)
    x[1, 5, 2] = np.nan
    got, status = noalloc(x, 2, 2, 0.0, "float32")
    assert status & 2
    x[1, 5, 2] = 0
    got, status = noalloc(x, 2, 2, 0.0, "float32")  # the bit does not linger in scratch
    assert status == 0
    lib = _native.lib()
    p = ctypes.c_void_p(1)
    # 1024 windows of 1024 x 16 float32 need more than the fixed scratch.
    assert lib.hdb_workspace_bytes(1024, 1024, 16, 0, 5, 0) > _native.NOALLOC_BYTES
    rc = lib.hdb_cluster_noalloc(None, p, 1024, 1024, 16, 0, 5, 5, 0.0, p, p, 0)
    assert rc != 0 and "scratch" in lib.hdb_last_error().decode()
