"""torch.ops.hdbscan_amd.cluster without a GPU: the fake (meta)
implementation's shapes and the argument checks, and tracing under
torch.compile(fullgraph=True).  Messages are the public function's
(reference tests/test_api.py:15-77)."""

import pytest

torch = pytest.importorskip("torch")
import hdbscan_amd.ops  # noqa: E402,F401  (registers the op)

OP = torch.ops.hdbscan_amd.cluster


@pytest.mark.parametrize("shape,out", [((3, 100, 8), (3, 100)), ((100, 8), (100,)), ((1, 2, 1), (1, 2))])
@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_fake_shapes(shape, out, precision):
    y = OP(torch.empty(shape, device="meta"), 2, 1, 0.0, precision)
    assert tuple(y.shape) == out and y.dtype == torch.int32 and y.device.type == "meta"


@pytest.mark.parametrize("shape,args,match", [
    ((10, 2), (5, 5, 0.0, "float16"), "precision"),
    ((10, 2), (1, 1, 0.0, "float32"), "min_cluster_size must be >= 2"),
    ((10, 2), (2, 0, 0.0, "float32"), "min_samples must be >= 1"),
    ((10, 2), (2, 1, -0.5, "float32"), "finite and >= 0"),
    ((10, 2), (2, 1, float("inf"), "float32"), "finite and >= 0"),
    ((10, 2), (2, 1, float("nan"), "float32"), "finite and >= 0"),
    ((10,), (2, 1, 0.0, "float32"), "rank 2"),
    ((1, 10, 2, 2), (2, 1, 0.0, "float32"), "rank 2"),
    ((1, 2), (2, 1, 0.0, "float32"), "at least 2 samples"),
    ((0, 5, 2), (2, 1, 0.0, "float32"), "nonempty"),
    ((10, 0), (2, 1, 0.0, "float32"), "nonempty"),
    ((4097, 2), (2, 1, 0.0, "float32"), "4096"),
    ((10, 2), (2, 11, 0.0, "float32"), "must not exceed"),
])
def test_checks(shape, args, match):
    with pytest.raises(ValueError, match=match):
        OP(torch.empty(shape, device="meta"), *args)


@pytest.mark.parametrize("dtype", [torch.bool, torch.complex64])
def test_non_real_rejected(dtype):
    with pytest.raises(ValueError, match="real numeric"):
        OP(torch.empty((10, 2), dtype=dtype, device="meta"), 2, 1, 0.0, "float32")


def test_traces_under_compile_fullgraph():
    def f(x):
        return OP(x * 2.0, 5, 5, 0.0, "float32") + 1

    g = torch.compile(f, backend="aot_eager", fullgraph=True)
    y = g(torch.empty(2, 64, 4, device="meta"))
    assert tuple(y.shape) == (2, 64) and y.dtype == torch.int32


def test_fake_tensor_mode():
    from torch._subclasses.fake_tensor import FakeTensorMode

    with FakeTensorMode():
        y = OP(torch.empty(2, 64, 4), 5, 5, 0.0, "float64")
        assert tuple(y.shape) == (2, 64) and y.dtype == torch.int32
