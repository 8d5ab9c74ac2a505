"""The MFMA-bounded matrix of the float32 matrix path (near pairs exact, far
pairs as certified lower bounds; Prim resolves a vertex exactly when a lower
bound could win) against the CPU oracle: Prim edges, every tree intermediate
and the labels bitwise, on windows that exercise its fallbacks (exact tiles
for norms past the bounded range, tiles with more near pairs than the lists
hold, overflowed core lists, resolution at cluster jumps).  Every call goes
through the C ABI."""

import numpy as np
import pytest

import gpu_helpers as G
from hdbscan_amd import _native
from hdbscan_amd._native import FLAG_NO_GRAM_EXACT, FLAG_PLAIN
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")


def _embed(rng, n, d, comps, spread, dup):
    means = rng.normal(size=(comps, d))
    owner = rng.integers(0, comps, n)
    pts = means[owner] + rng.normal(0, spread, (n, d))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    k = int(round(n * dup))
    if k:
        pts[rng.choice(n, k, replace=False)] = pts[rng.integers(0, n, k)]
    return pts


def _windows():
    rng = np.random.default_rng(20261204)
    base = rng.normal(size=(320, 200))
    out = {
        "embed-300x128": _embed(rng, 300, 128, 6, 0.3, 0.03),
        "embed-1000x100": _embed(rng, 1000, 100, 9, 0.4, 0.05),
        "embed-2048x128": _embed(rng, 2048, 128, 16, 0.35, 0.03),
        "embed-4096x96": _embed(rng, 4096, 96, 24, 0.4, 0.02),
        "blobs-517x40": np.concatenate([rng.normal(c, 1.0, (47, 40)) for c in rng.uniform(-6, 6, (11, 40))]),
        "blobs-700x200": np.concatenate([rng.normal(c, 0.7, (70, 200)) for c in rng.uniform(-3, 3, (10, 200))]),
        "grid-ties-400x70": rng.integers(-2, 3, (400, 70)) / 4.0,
        "offset-1e3-256x96": base[:256, :96] + 1e3,
        "duplicates-320x80": np.repeat(base[:40, :80], 8, axis=0),
        "all-equal-200x65": np.ones((200, 65)),
        # Every pair near: the list overflows and interior tiles hold near
        # pairs evaluated by the matrix kernel itself.
        "all-equal-300x64": np.full((300, 64), 3.0),
        "clumps-512x32": base[:4, :32][np.arange(512) % 4] + 0.0,
        "huge-in-256x72": base[:256, :72] * 2.0 ** 40,
        "huge-out-256x72": base[:256, :72] * 2.0 ** 52,
        "tiny-256x72": base[:256, :72] * 2.0 ** -60,
        "two-2x100": base[:2, :100],
        "three-3x90": base[:3, :90],
        "small-64x128": base[:64, :128],
        "d33-300x33": base[:300, :33],
        "d32-260x32": base[:260, :32],
        "d20-777x20": np.concatenate([rng.normal(c, 0.5, (111, 20)) for c in rng.uniform(-4, 4, (7, 20))]),
    }
    return {k: v.astype(np.float32) for k, v in out.items()}


WINDOWS = _windows()
CASES = [(name, ms) for name in WINDOWS for ms in (1, 2, 3, 5, 8, 16) if ms <= len(WINDOWS[name])]


def _stage_names(x, mcs, ms, flags=0):
    lib = _native.lib()
    lib.hdb_profile(1)
    try:
        G.cluster(x, mcs, ms, 0.0, "float32", flags=flags)
        names = [name for name, _ in _native.stage_times()]
    finally:
        lib.hdb_profile(0)  # also clears the recorded stages
    return names


def _oracle(x, mcs, ms):
    dist = oracle.distances(x, "float32")
    core = oracle.core(dist, ms, "float32") if ms > 2 else None
    u, v, w = oracle.prim(dist, core, unit_core=ms <= 2, precision="float32")
    return (u, v, w), oracle.tree(u, v, w, mcs, 0.0, precision="float32")


@pytest.mark.parametrize("name,ms", CASES, ids=[f"{n}-ms{m}" for n, m in CASES])
def test_bounded_matrix_matches_oracle(name, ms):
    """The float32 matrix path (HDB_FLAG_MATRIX) takes the MFMA-bounded
    matrix for min_samples <= 16; edges, tree and labels equal the oracle's
    bit for bit."""
    x = WINDOWS[name]
    mcs = 2 if len(x) < 10 else 5
    flags = _native.FLAG_MATRIX
    assert "gx_matrix_kernel" in _stage_names(x, mcs, ms, flags)
    labels, status, ex = G.cluster_export(x, mcs, ms, 0.0, "float32", flags=flags)
    assert status == 0
    (u, v, w), t = _oracle(x, mcs, ms)
    np.testing.assert_array_equal(ex[0]["edge_u"], u)
    np.testing.assert_array_equal(ex[0]["edge_v"], v)
    np.testing.assert_array_equal(ex[0]["edge_w"].view(np.uint32), w.view(np.uint32))
    for k in ("order", "left", "right", "sizes", "parent", "selected", "point_parent"):
        np.testing.assert_array_equal(ex[0][k], t[k], err_msg=k)
    for k in ("birth", "stability"):
        np.testing.assert_array_equal(ex[0][k].view(np.uint32), t[k].view(np.uint32), err_msg=k)
    np.testing.assert_array_equal(labels[0], t["labels"])


def test_default_plan_takes_bounded_matrix_and_flags_turn_it_off():
    """D=128 float32 windows take the bounded matrix by default;
    HDB_FLAG_NO_GRAM_EXACT and HDB_FLAG_PLAIN give the plain exact matrix
    (distance_kernel), with the same labels."""
    x = WINDOWS["embed-300x128"]
    assert "gx_matrix_kernel" in _stage_names(x, 5, 8)
    for flags in (FLAG_NO_GRAM_EXACT, FLAG_PLAIN):
        names = _stage_names(x, 5, 8, flags)
        assert "distance_kernel" in names and "gx_matrix_kernel" not in names
    a, sa = G.cluster(x, 5, 8, 0.0, "float32")
    b, sb = G.cluster(x, 5, 8, 0.0, "float32", flags=FLAG_NO_GRAM_EXACT)
    assert sa == sb == 0
    np.testing.assert_array_equal(a, b)


def test_bounded_matrix_batch_of_windows():
    """Several windows in one call (per-window matrices, lists and Prim)."""
    rng = np.random.default_rng(5)
    x = np.stack([_embed(rng, 600, 128, 5 + b, 0.35, 0.03) for b in range(6)]).astype(np.float32)
    labels, status = G.cluster(x, 10, 6, 0.0, "float32")
    assert status == 0
    for b in range(len(x)):
        expected, st = oracle.hdbscan(x[b], 10, 6, 0.0, precision="float32")
        assert st == 0
        np.testing.assert_array_equal(labels[b], expected, err_msg=f"window {b}")


@pytest.mark.parametrize("name", [k for k in WINDOWS if len(WINDOWS[k]) >= 8])
@pytest.mark.parametrize("ms", [3, 8, 16])
def test_bounded_matrix_entries_are_exact_or_certified(name, ms):
    """Every entry of the bounded matrix is either the oracle's exact
    distance, bit for bit (sign bit clear), or minus a lower bound of it
    (sign bit set); the diagonal is +0 and both halves agree.  Cancellation
    (offset windows), norms near and past the bounded range, ties and
    duplicates exercise the MFMA certificate."""
    x = WINDOWS[name]
    got, status = G.bounded_distances(x, ms)
    assert status == 0
    m = got[0]
    exact = oracle.distances(x, "float32")
    far = np.signbit(m)
    np.testing.assert_array_equal(m[~far].view(np.uint32), exact[~far].view(np.uint32))
    assert (-m[far] <= exact[far]).all(), float((-m[far] - exact[far]).max())
    np.testing.assert_array_equal(m.view(np.uint32), m.T.view(np.uint32))
    assert (np.diag(m).view(np.uint32) == 0).all()


def test_bounded_matrix_entries_batch_and_odd_sizes():
    """hdb_bounded_distances over a batch of windows (per-window matrices
    and lists) whose size is not a multiple of 4 (no 16-byte mirror stores)
    and spans partial tiles: every entry exact or a certified lower bound."""
    rng = np.random.default_rng(11)
    x = np.stack([_embed(rng, 389, 96, 4 + b, 0.3, 0.04) for b in range(5)]).astype(np.float32)
    got, status = G.bounded_distances(x, 6)
    assert status == 0
    for b in range(len(x)):
        m = got[b]
        exact = oracle.distances(x[b], "float32")
        far = np.signbit(m)
        assert (~far).sum() > len(m)  # near pairs beyond the diagonal
        np.testing.assert_array_equal(m[~far].view(np.uint32), exact[~far].view(np.uint32), err_msg=f"window {b}")
        assert (-m[far] <= exact[far]).all(), b
        np.testing.assert_array_equal(m.view(np.uint32), m.T.view(np.uint32))


@pytest.mark.parametrize("value,bit", [(np.nan, 2), (np.inf, 2), (-np.inf, 2), (1e20, 1)])
def test_bounded_matrix_status(value, bit):
    """Non-finite input (bit 2, from the prep kernel) and a distance that
    overflows float32 (bit 1, from an exact tile: the point's squared norm is
    beyond the bounded range)."""
    x = WINDOWS["embed-300x128"].copy()
    x[17, 5] = value
    _, status = G.cluster(x, 5, 8, 0.0, "float32")
    assert status & bit
    if bit == 2:
        assert status == 2 or status == 3
