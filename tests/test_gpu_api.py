"""Public API on the GPU: the reference's API, float64 and parity tests
(reference tests/test_api.py:86-287, tests/test_float64.py, tests/test_parity.py)
against hdbscan_amd.hdbscan(backend="rocm")."""

import os
import subprocess
import sys

import numpy as np
import pytest

from golden_io import same_partition
from oracle import oracle

pytestmark = pytest.mark.gpu

VALUES = np.array([[0], [0.01], [0.02], [3], [3.01], [3.02]], np.float32)


@pytest.fixture(scope="module")
def torch():
    t = pytest.importorskip("torch")
    if not t.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return t


def run(x, **kw):
    from hdbscan_amd import hdbscan

    return hdbscan(x, backend="rocm", **kw)


def host(labels, torch):
    torch.cuda.synchronize()
    return labels.cpu().numpy()


def expect_labels(labels, shape, torch):
    assert isinstance(labels, torch.Tensor)
    assert labels.is_cuda and labels.dtype == torch.int32
    assert tuple(labels.shape) == shape
    assert labels.__dlpack_device__()[0] == 10  # kDLROCM
    out = host(labels, torch)
    assert np.all(out >= -1)
    return out


def oracle_labels(x, mcs, ms=None, eps=0.0, precision="float32"):
    labels, status = oracle.hdbscan(np.asarray(x), mcs, ms, eps, precision=precision)
    assert status == 0
    return labels


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("dtype", [np.float16, np.float32, np.float64, np.int32, np.uint32])
def test_host_numeric_dtypes_and_no_mutation(torch, precision, dtype):
    values = (VALUES * 100).astype(dtype)
    before = values.copy()
    values.flags.writeable = False
    labels = run(values, min_cluster_size=2, min_samples=1, precision=precision)
    got = expect_labels(labels, (6,), torch)
    np.testing.assert_array_equal(got, oracle_labels(values.astype(precision), 2, 1, precision=precision))
    np.testing.assert_array_equal(values, before)


@pytest.mark.parametrize("layout", ["reversed", "column-strided", "fortran", "list"])
def test_host_strides_lists_and_numpy_scalars(torch, layout):
    values = np.concatenate([VALUES, VALUES + 1, VALUES + 2, VALUES + 3], axis=1)
    if layout == "reversed":
        values = values[::-1, ::-1]
    elif layout == "column-strided":
        values = values[:, ::2]
    elif layout == "fortran":
        values = np.asfortranarray(values)
    expected = oracle_labels(np.ascontiguousarray(values, np.float32), 2, 1)
    labels = run(values.tolist() if layout == "list" else values, min_cluster_size=np.int64(2),
                 min_samples=np.int32(1), cluster_selection_epsilon=np.float64(0))
    np.testing.assert_array_equal(expect_labels(labels, (6,), torch), expected)


def test_native_strided_input_and_no_mutation(torch):
    host_values = np.arange(48, dtype=np.float32).reshape(8, 6)
    native = torch.from_numpy(host_values).to("cuda")
    views = [native[:, ::2], native.flip(0)[:, 1::2], native.T,
             torch.from_numpy(np.stack([host_values, host_values + 100])).to("cuda")[1:]]
    for view in views:
        before = view.cpu().numpy().copy()
        labels = run(view, min_cluster_size=2, min_samples=1)
        got = expect_labels(labels, before.shape[:-1], torch)
        np.testing.assert_array_equal(view.cpu().numpy(), before)
        if before.ndim == 2:
            np.testing.assert_array_equal(got, oracle_labels(before, 2, 1))
        else:
            np.testing.assert_array_equal(got[0], oracle_labels(before[0], 2, 1))
    np.testing.assert_array_equal(native.cpu().numpy(), host_values)


@pytest.mark.parametrize("value", [float("nan"), float("inf"), -float("inf")])
def test_native_nonfinite_rejected(torch, value):
    values = torch.full((6, 2), value, device="cuda")
    with pytest.raises(ValueError, match="finite"):
        run(values)


@pytest.mark.parametrize("dtype", ["bool", "complex64"])
def test_native_nonnumeric_rejected(torch, dtype):
    values = torch.from_numpy(VALUES).to("cuda").to(getattr(torch, dtype))
    with pytest.raises(ValueError, match="real numeric"):
        run(values)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_overflow_rejected(torch, precision):
    largest = np.finfo(precision).max / 4
    values = np.array([[-largest], [-largest], [largest], [largest]])
    with pytest.raises(ValueError, match="overflow"):
        run(values, min_cluster_size=2, precision=precision)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_requested_precision_conversion_rejects_nonfinite(torch, precision):
    if precision == "float64":
        values = torch.full((6, 1), float("inf"), dtype=torch.float64, device="cuda")
    else:
        values = torch.full((6, 1), 1e40, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="finite"):
        run(values, precision=precision)


def test_nonfinite_outranks_overflow_in_a_partial_wave(torch):
    values = np.zeros((3, 33, 1), dtype=np.float32)
    values[-1, -1, 0] = np.finfo(np.float32).max
    with pytest.raises(ValueError, match="overflow"):
        run(torch.from_numpy(values).to("cuda"), min_cluster_size=2)
    values[0, 0, 0] = np.finfo(np.float32).max
    values[-1, -1, 0] = np.nan
    with pytest.raises(ValueError, match="only finite"):
        run(torch.from_numpy(values).to("cuda"), min_cluster_size=2)


def test_nondefault_stream_is_used_and_preserved(torch):
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        labels = run(VALUES, min_cluster_size=2, min_samples=1)
        assert torch.cuda.current_stream() == stream
    stream.synchronize()
    np.testing.assert_array_equal(expect_labels(labels, (6,), torch), oracle_labels(VALUES, 2, 1))


GAP = 2.0 ** -30
COLLAPSING = np.repeat(1 + np.array([0, GAP, 16 * GAP, 17 * GAP], dtype=np.float64), 2)[:, None]
BOUNDARY_POINTS = np.repeat(1 + 2.0 ** -40 + np.array([0, 0.125, 4, 4.125], dtype=np.float64), 2)[:, None]
BOUNDARY = np.array([[0], [0], [0.125], [0.125], [4], [4], [4.125], [4.125]], dtype=np.float32)


def test_float64_preserves_coordinates(torch):
    for values in (COLLAPSING, torch.from_numpy(COLLAPSING).to("cuda")):
        labels = run(values, precision="float64", min_cluster_size=2, min_samples=1)
        assert same_partition(expect_labels(labels, (8,), torch), np.repeat(np.arange(4), 2))


def test_float64_epsilon_boundary(torch):
    windows = np.stack([BOUNDARY_POINTS, BOUNDARY_POINTS[::-1]])
    native = torch.from_numpy(windows).to("cuda")
    for eps, clusters in ((float(np.nextafter(0.125, 0)), 4), (0.125, 4), (float(np.nextafter(0.125, 1)), 2)):
        labels = expect_labels(run(native, precision="float64", cluster_selection_epsilon=eps,
                                   min_cluster_size=2, min_samples=1), (2, 8), torch)
        for got, window in zip(labels, windows):
            np.testing.assert_array_equal(got, oracle_labels(window, 2, 1, eps, "float64"))
        assert same_partition(labels[0], np.repeat(np.arange(clusters), 8 // clusters))


@pytest.mark.parametrize("eps,clusters", [
    (0.0, 4), (0.1, 4), (float(np.nextafter(np.float32(0.125), np.float32(0))), 4), (0.125, 4),
    (float(np.nextafter(np.float32(0.125), np.float32(1))), 2)])
def test_epsilon_boundary_float32(torch, eps, clusters):
    labels = expect_labels(run(BOUNDARY, min_cluster_size=2, min_samples=1, cluster_selection_epsilon=eps),
                           (8,), torch)
    assert same_partition(labels, np.repeat(np.arange(clusters), 8 // clusters))


def test_documented_float32_epsilon_resolution_limit(torch):
    eps = float(np.nextafter(np.float64(0.125), np.float64(1)))
    got32 = expect_labels(run(BOUNDARY, cluster_selection_epsilon=eps, min_cluster_size=2, min_samples=1),
                          (8,), torch)
    got64 = expect_labels(run(BOUNDARY, cluster_selection_epsilon=eps, min_cluster_size=2, min_samples=1,
                              precision="float64"), (8,), torch)
    assert same_partition(got32, [0, 0, 1, 1, 2, 2, 3, 3])
    assert same_partition(got64, [0, 0, 0, 0, 1, 1, 1, 1])


@pytest.mark.parametrize("epsilon", [0.0, 0.1])
def test_batch_individual_and_position_invariance(torch, epsilon):
    rng = np.random.default_rng(31415)
    windows = np.stack([rng.normal(size=(32, 3)), rng.integers(-4, 5, (32, 3)) / 2,
                        np.zeros((32, 3))]).astype(np.float32)
    p = dict(min_cluster_size=3, min_samples=2, cluster_selection_epsilon=epsilon)
    individual = [expect_labels(run(w, **p), (32,), torch) for w in windows]
    for w, got in zip(windows, individual):
        np.testing.assert_array_equal(got, oracle_labels(w, 3, 2, epsilon))
    for order in ([0, 1, 2], [2, 0, 1], [1, 2, 0], [0, 0, 2]):
        got = expect_labels(run(windows[order], **p), (3, 32), torch)
        for position, labels in enumerate(got):
            np.testing.assert_array_equal(labels, individual[order[position]])
    single = expect_labels(run(windows[:1], **p), (1, 32), torch)
    np.testing.assert_array_equal(single[0], individual[0])


def test_default_min_samples(torch):
    rng = np.random.default_rng(20261010)
    values = np.concatenate([rng.normal(c, 0.08, (24, 2)) for c in ((-3, -3), (0, 3), (3, -3))]).astype(np.float32)
    implicit = expect_labels(run(values, min_cluster_size=4), (72,), torch)
    explicit = expect_labels(run(values, min_cluster_size=4, min_samples=4), (72,), torch)
    np.testing.assert_array_equal(implicit, explicit)
    assert same_partition(implicit, np.repeat(np.arange(3), 24))


def test_production_call_without_sklearn_or_oracle_or_host_copy(torch):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = r"""
import importlib.abc, sys
class Block(importlib.abc.MetaPathFinder):
    def find_spec(self, name, path=None, target=None):
        if name.split('.')[0] in ('sklearn', 'oracle', 'scipy'):
            raise ImportError('blocked ' + name)
        return None
sys.meta_path.insert(0, Block())
import numpy as np, torch
import hdbscan_amd
real_asarray = np.asarray
def guarded(a, *args, **kwargs):
    if isinstance(a, torch.Tensor) and a.is_cuda:
        raise AssertionError('device input copied to host')
    return real_asarray(a, *args, **kwargs)
np.asarray = guarded
x = torch.arange(64, dtype=torch.float32, device='cuda').reshape(32, 2) % 7
labels = hdbscan_amd.hdbscan(x, min_cluster_size=2)
torch.cuda.synchronize()
assert labels.is_cuda and labels.dtype == torch.int32
print('ok')
"""
    out = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, cwd=root, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip().endswith("ok")


def test_custom_op_eager_and_compiled(torch):
    """torch.ops.hdbscan_amd.cluster on device tensors: eager, and inside a
    torch.compile(fullgraph=True) graph, label-exact against the oracle."""
    import hdbscan_amd.ops  # noqa: F401

    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(size=(3, 120, 4)), rng.normal(size=(3, 80, 4)) + 6], axis=1)
    x = x.astype(np.float32)
    xd = torch.from_numpy(x).to("cuda")
    eager = torch.ops.hdbscan_amd.cluster(xd, 5, 5, 0.0, "float32")
    single = torch.ops.hdbscan_amd.cluster(xd[1], 5, 5, 0.0, "float32")

    def f(t):
        return torch.ops.hdbscan_amd.cluster(t * 1.0, 5, 5, 0.0, "float32")

    compiled = torch.compile(f, backend="aot_eager", fullgraph=True)(xd)
    out = expect_labels(eager, (3, 200), torch)
    np.testing.assert_array_equal(expect_labels(compiled, (3, 200), torch), out)
    np.testing.assert_array_equal(expect_labels(single, (200,), torch), out[1])
    for b in range(3):
        np.testing.assert_array_equal(out[b], oracle_labels(x[b], 5, 5))


def test_custom_op_reports_nonfinite(torch):
    import hdbscan_amd.ops  # noqa: F401

    x = torch.zeros((1, 16, 2), device="cuda")
    x[0, 3, 1] = float("nan")
    with pytest.raises(ValueError, match="finite"):
        torch.ops.hdbscan_amd.cluster(x, 2, 2, 0.0, "float32")


# Shapes that take the materialised-matrix path (float32 beyond 64
# dimensions; float64 with 32 dimensions above 1024 points): the input check
# must report a non-finite coordinate as such there too, and it must outrank
# the overflow that the same coordinate causes in the distances
# (reference tests/test_api.py:140-174, tests/test_kernels.py:703-715).
MATRIX_SHAPES = [("float32", 300, 128), ("float64", 2048, 32)]


# 1e40 is finite in float64 and +inf once converted to float32.
NONFINITE_CASES = [(p, n, d, v) for p, n, d in MATRIX_SHAPES
                   for v in [float("nan"), float("inf"), -float("inf")] + (["1e40"] if p == "float32" else [])]


@pytest.mark.parametrize("precision,n,d,value", NONFINITE_CASES)
@pytest.mark.parametrize("entry", ["hdbscan", "hdb_cluster"])
def test_matrix_path_nonfinite_reported_as_input(torch, precision, n, d, value, entry):
    import gpu_helpers as G

    rng = np.random.default_rng(n + d)
    x = rng.normal(size=(2, n, d))
    # An overflowing pair in the first window, the non-finite coordinate in
    # the last row of the second (the last tile and the last chunk of
    # dimensions).
    x[0, 0, 0] = -np.finfo(precision).max / 4
    x[0, 1, 0] = np.finfo(precision).max / 4
    if value == "1e40":
        x[1, -1, -1] = 1e40
    else:
        x[1, -1, -1] = value
    if entry == "hdbscan":
        xd = torch.from_numpy(x).to("cuda")  # float64 on the device, converted by the backend
        with pytest.raises(ValueError, match="only finite values in the requested precision"):
            run(xd, min_cluster_size=5, min_samples=5, precision=precision)
    else:
        with np.errstate(over="ignore"):
            xr = x.astype(precision)
        _, status = G.cluster(xr, 5, 5, 0.0, precision)
        assert status & 2, status
        # The first window alone overflows only.
        _, status = G.cluster(xr[:1], 5, 5, 0.0, precision)
        assert status == 1


@pytest.mark.parametrize("precision,n,d", MATRIX_SHAPES)
def test_matrix_path_overflow_alone(torch, precision, n, d):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(n, d)).astype(precision)
    x[0, 0] = -np.finfo(precision).max / 4
    x[1, 0] = np.finfo(precision).max / 4
    with pytest.raises(ValueError, match="overflow"):
        run(torch.from_numpy(x).to("cuda"), min_cluster_size=5, min_samples=5, precision=precision)
