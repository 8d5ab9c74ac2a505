"""ctypes binding of libhdbscan_rocm.so (declared in include/hdbscan_rocm.h)."""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "_lib", "libhdbscan_rocm.so")

F32, F64 = 0, 1
STATUS_OVERFLOW, STATUS_NONFINITE = 1, 2
FLAG_GLOBAL_ARENA, FLAG_MATRIX, FLAG_NO_POINT_LDS, FLAG_MATRIX_INPUT, FLAG_SERIAL_SORT = 1, 2, 4, 8, 16
FLAG_PREFILTER = 32
FLAG_PLAIN = 64
FLAG_FILTER = 128
FLAG_ONE_STEP = 65536
FLAG_GRAM_CHUNKED = 131072
FLAG_GRAM_GRAPH = 262144
FLAG_NO_GRAPH = 524288
FLAG_SPLIT_TREE = 1048576
FLAG_FUSED_TREE = 2097152
FLAG_SERIAL_LINKAGE = 4194304
FLAG_NO_GRAM_EXACT = 8388608
EXPORTED = ("hdb_version", "hdb_last_error", "hdb_profile", "hdb_stage_count", "hdb_stage_name",
            "hdb_stage_ms", "hdb_workspace_bytes", "hdb_cluster", "hdb_cluster_export", "hdb_distances",
            "hdb_core_distances", "hdb_cluster_matrix", "hdb_ragged_workspace_bytes", "hdb_cluster_ragged",
            "hdb_cluster_noalloc", "hdb_probe_valu", "hdb_ragged_set_tuning", "hdb_bounded_distances")
NOALLOC_BYTES = 64 * 1024 * 1024

_lib = None
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_I = ctypes.c_int


def lib():
    """Load the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise ImportError(
                f"hdbscan_amd native library missing ({LIBRARY}); build it with "
                "python -m hdbscan_amd.build")
        handle = ctypes.CDLL(LIBRARY)
        handle.hdb_version.restype = _I
        handle.hdb_version.argtypes = []
        handle.hdb_last_error.restype = ctypes.c_char_p
        handle.hdb_last_error.argtypes = []
        handle.hdb_profile.restype = _I
        handle.hdb_profile.argtypes = [_I]
        handle.hdb_stage_count.restype = _I
        handle.hdb_stage_count.argtypes = []
        handle.hdb_stage_name.restype = ctypes.c_char_p
        handle.hdb_stage_name.argtypes = [_I]
        handle.hdb_stage_ms.restype = ctypes.c_float
        handle.hdb_stage_ms.argtypes = [_I]
        handle.hdb_ragged_set_tuning.restype = _I
        handle.hdb_ragged_set_tuning.argtypes = [_P, _I]
        handle.hdb_probe_valu.restype = _I
        handle.hdb_probe_valu.argtypes = [_P, _I, _I, _P, _P]
        handle.hdb_workspace_bytes.restype = ctypes.c_size_t
        handle.hdb_workspace_bytes.argtypes = [_I64, _I64, _I64, _I, _I32, _I]
        handle.hdb_cluster.restype = _I
        handle.hdb_cluster.argtypes = [_P, _P, _I64, _I64, _I64, _I, _I64, _I32, ctypes.c_double,
                                       _P, _P, ctypes.c_size_t, _P, _I]
        handle.hdb_cluster_noalloc.restype = _I
        handle.hdb_cluster_noalloc.argtypes = [_P, _P, _I64, _I64, _I64, _I, _I64, _I32, ctypes.c_double,
                                               _P, _P, _I]
        handle.hdb_cluster_export.restype = _I
        handle.hdb_cluster_export.argtypes = [_P, _P, _I64, _I64, _I64, _I, _I64, _I32, ctypes.c_double, _I,
                                              _P, _P, ctypes.c_size_t, _P, _I, _P, _P]
        handle.hdb_distances.restype = _I
        handle.hdb_distances.argtypes = [_P, _P, _I64, _I64, _I64, _I, _P, _P]
        handle.hdb_bounded_distances.restype = _I
        handle.hdb_bounded_distances.argtypes = [_P, _P, _I64, _I64, _I64, _I, _I32, _P, _P, ctypes.c_size_t, _P]
        handle.hdb_core_distances.restype = _I
        handle.hdb_core_distances.argtypes = [_P, _P, _I64, _I64, _I64, _I, _I32, _P, _P,
                                              ctypes.c_size_t, _P, _I]
        handle.hdb_cluster_matrix.restype = _I
        handle.hdb_cluster_matrix.argtypes = [_P, _P, _P, _I64, _I64, _I, _I64, _I, ctypes.c_double,
                                              _I, _P, _P, ctypes.c_size_t, _P, _I, _P, _P]
        handle.hdb_ragged_workspace_bytes.restype = ctypes.c_size_t
        handle.hdb_ragged_workspace_bytes.argtypes = [_P, _I64, _I64, _I, _I32, _I]
        handle.hdb_cluster_ragged.restype = _I
        handle.hdb_cluster_ragged.argtypes = [_P, _P, _P, _P, _I64, _I64, _I, _I64, _I32, ctypes.c_double,
                                              _P, _P, ctypes.c_size_t, _P, _I]
        _lib = handle
    return _lib


def stage_times():
    """[(kernel name, ms)] of the last profiled call on this thread."""
    handle = lib()
    return [(handle.hdb_stage_name(i).decode(), float(handle.hdb_stage_ms(i)))
            for i in range(handle.hdb_stage_count())]


def check(rc):
    if rc != 0:
        raise RuntimeError("hdbscan_rocm: " + lib().hdb_last_error().decode())
