"""Public API contract (CPU): the reference's validation cases
(reference tests/test_api.py:15-83) against hdbscan_amd.hdbscan, which must
reject them before anything touches a device."""

import builtins

import numpy as np
import pytest

from hdbscan_amd import hdbscan

VALUES = np.array([[0], [0.01], [0.02], [3], [3.01], [3.02]], np.float32)


@pytest.mark.parametrize("parameters,match", [
    ({"backend": "cpu"}, "backend"),
    ({"backend": "auto"}, "backend"),
    ({"backend": "ROCM"}, "backend"),
    ({"backend": "cuda"}, "backend"),
    ({"backend": None}, "backend"),
    ({"precision": "float16"}, "precision"),
    ({"precision": np.float32}, "precision"),
    ({"min_cluster_size": 1}, "min_cluster_size"),
    ({"min_cluster_size": 2.0}, "min_cluster_size"),
    ({"min_cluster_size": True}, "min_cluster_size"),
    ({"min_cluster_size": np.bool_(True)}, "min_cluster_size"),
    ({"min_cluster_size": None}, "min_cluster_size"),
    ({"min_cluster_size": 2**31, "min_samples": 1}, "int32"),
    ({"min_samples": 0}, "min_samples"),
    ({"min_samples": -1}, "min_samples"),
    ({"min_samples": 2.0}, "min_samples"),
    ({"min_samples": True}, "min_samples"),
    ({"min_samples": np.bool_(True)}, "min_samples"),
    ({"min_samples": 7}, "min_samples"),
    ({"cluster_selection_epsilon": -0.1}, "epsilon"),
    ({"cluster_selection_epsilon": float("nan")}, "epsilon"),
    ({"cluster_selection_epsilon": float("inf")}, "epsilon"),
    ({"cluster_selection_epsilon": True}, "epsilon"),
    ({"cluster_selection_epsilon": "0.1"}, "epsilon"),
    ({"cluster_selection_epsilon": 1e40}, "representable"),
])
def test_parameter_validation_without_device_dispatch(monkeypatch, parameters, match):
    original_import = builtins.__import__

    def no_backend_import(name, *args, **kwargs):
        if name in ("_rocm", "torch", "hdbscan_amd._rocm"):
            raise AssertionError(f"Invalid parameters dispatched to {name}")
        return original_import(name, *args, **kwargs)

    monkeypatch.setattr(builtins, "__import__", no_backend_import)
    with pytest.raises(ValueError, match=match):
        hdbscan(VALUES, **parameters)


@pytest.mark.parametrize("values,match", [
    (np.zeros(6), "rank"),
    (np.array(0), "rank"),
    (np.zeros((1, 2, 3, 4)), "rank"),
    (np.zeros((0, 2)), "at least 2"),
    (np.zeros((1, 2)), "at least 2"),
    (np.zeros((2, 0)), "nonempty"),
    (np.zeros((0, 6, 2)), "nonempty"),
    (np.zeros((6, 2), dtype=bool), "real numeric"),
    (np.ones((6, 2), dtype=complex), "real numeric"),
    (np.full((6, 2), "1"), "real numeric"),
    (np.full((6, 2), object(), dtype=object), "real numeric"),
    (np.full((6, 2), np.nan), "finite"),
    (np.full((6, 2), np.inf), "finite"),
    (np.full((6, 2), -np.inf), "finite"),
    (np.zeros((4097, 1), np.float32), "4096"),
    (np.broadcast_to(np.float32(0), (5, 4096, 1)), "512 MiB"),
])
def test_host_input_validation(monkeypatch, values, match):
    original_import = builtins.__import__

    def no_backend_import(name, *args, **kwargs):
        if name in ("_rocm", "torch"):
            raise AssertionError(f"Invalid input dispatched to {name}")
        return original_import(name, *args, **kwargs)

    monkeypatch.setattr(builtins, "__import__", no_backend_import)
    with pytest.raises(ValueError, match=match):
        hdbscan(values, min_cluster_size=2, min_samples=1)


def test_float64_workspace_limit_before_dispatch():
    values = np.broadcast_to(np.float32(0), (3, 4096, 1))
    with pytest.raises(ValueError, match="512 MiB"):
        hdbscan(values, precision="float64")


def test_import_does_not_load_torch_or_the_library():
    import subprocess
    import sys

    code = ("import sys; import hdbscan_amd; "
            "assert 'torch' not in sys.modules; assert 'hdbscan_amd._rocm' not in sys.modules; print('ok')")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


class _ForeignDeviceArray:
    """Stands for another framework's device array (CuPy, JAX, Numba ...)."""

    def __init__(self, protocol):
        if protocol == "cuda":
            self.__cuda_array_interface__ = {"shape": (6, 1), "typestr": "<f4", "data": (0, False),
                                             "version": 3}
        else:
            self.__dlpack_device__ = lambda: (2, 0)

    def __array__(self, *args, **kwargs):
        raise AssertionError("a foreign device array must not be read through the host")


@pytest.mark.parametrize("protocol", ["cuda", "dlpack"])
def test_foreign_device_arrays_rejected_without_host_copy(monkeypatch, protocol):
    """reference tests: a non-native object exposing a device protocol raises
    TypeError before any conversion (reference hdbscan_gpu/_api.py:62-64)."""
    original_import = builtins.__import__

    def no_backend_import(name, *args, **kwargs):
        if name in ("_rocm", "hdbscan_amd._rocm"):
            raise AssertionError(f"Foreign input dispatched to {name}")
        return original_import(name, *args, **kwargs)

    monkeypatch.setattr(builtins, "__import__", no_backend_import)
    with pytest.raises(TypeError, match="native ROCm torch tensor or a NumPy host array"):
        hdbscan(_ForeignDeviceArray(protocol), min_cluster_size=2, min_samples=1)


def test_numpy_arrays_with_dlpack_are_host_data():
    """NumPy arrays expose __dlpack_device__ too; they stay host input."""
    assert hasattr(np.zeros(1), "__dlpack_device__")
    with pytest.raises(ValueError, match="finite"):
        hdbscan(np.full((6, 2), np.nan), min_cluster_size=2, min_samples=1)
