"""The C ABI library (CPU, no GPU calls): it loads, exports every function
that include/hdbscan_rocm.h declares, and rejects invalid arguments before
touching a device."""

import ctypes
import os
import re
import subprocess

import pytest

from hdbscan_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hdbscan_rocm.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hdb_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    for name in ("hdb_cluster", "hdb_workspace_bytes", "hdb_distances", "hdb_core_distances",
                 "hdb_cluster_matrix", "hdb_version", "hdb_last_error"):
        assert name in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIBRARY], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s(\w+)", out.stdout))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    lib = _native.lib()
    for name in declared():
        getattr(lib, name)
    assert set(_native.EXPORTED) == set(declared())


def test_version_and_workspace_sizing():
    lib = _native.lib()
    assert lib.hdb_version() >= 100
    ws = lib.hdb_workspace_bytes(256, 1024, 16, _native.F32, 5, _native.FLAG_PLAIN)
    assert ws > 256 * 1024 * 12
    # The MFMA filter adds per point: candidate lists (32 x 2 B) and counts
    # (4 B); its chunked form also bf16 pieces (2 x 2 x 16 x 2 B) and norm
    # terms (12 B), which the per-window form computes in LDS.
    wf = lib.hdb_workspace_bytes(256, 1024, 16, _native.F32, 5, 0)
    assert wf >= ws + 256 * 1024 * (64 + 4)
    wc = lib.hdb_workspace_bytes(256, 1024, 16, _native.F32, 5, _native.FLAG_GRAM_CHUNKED)
    assert wc >= ws + 256 * 1024 * (128 + 12 + 64 + 4)
    # The points path keeps 16 B per point for Prim's repacking.
    assert ws >= 256 * 1024 * (12 + 16)
    # The materialised matrix path needs B*N*N reals (and no repacking).
    wm = lib.hdb_workspace_bytes(256, 1024, 16, _native.F32, 5, _native.FLAG_MATRIX)
    assert wm >= 256 * 1024 * 1024 * 4 + 256 * 1024 * 12
    assert lib.hdb_workspace_bytes(4, 1024, 16, _native.F32, 100, 0) >= 4 * 1024 * 1024 * 4
    # 8 vertices per thread: accepted on the speculative matrix path.
    assert lib.hdb_workspace_bytes(4, 4096, 16, _native.F64, 5, _native.FLAG_MATRIX | 8 << 8) > 0


@pytest.mark.parametrize("args,match", [
    ((1, 1, 1, 0, 1, 0), "at least 2"),
    ((0, 8, 1, 0, 1, 0), "nonempty"),
    ((1, 4097, 1, 0, 1, 0), "4096"),
    ((1, 8, 1, 7, 1, 0), "dtype"),
    ((1, 8, 1, 0, 9, 0), "min_samples"),
    ((1, 8, 1, 0, 2, 3 << 8), "vertices per thread"),
    ((1, 8, 1, 0, 2, 8 << 8), "speculative matrix path"),
    ((1, 8, 1, 0, 2, 2 | 65536 | 8 << 8), "speculative matrix path"),
])
def test_invalid_sizes_report_errors(args, match):
    lib = _native.lib()
    assert lib.hdb_workspace_bytes(*args) == 0
    assert match in lib.hdb_last_error().decode()


def test_cluster_rejects_bad_arguments_without_a_device():
    lib = _native.lib()
    p = ctypes.c_void_p(16)
    rc = lib.hdb_cluster(None, p, 1, 8, 2, 0, 1, 2, 0.0, p, p, 1 << 20, p, 0)
    assert rc != 0 and "min_cluster_size" in lib.hdb_last_error().decode()
    rc = lib.hdb_cluster(None, None, 1, 8, 2, 0, 2, 2, 0.0, p, p, 1 << 20, p, 0)
    assert rc != 0 and "null" in lib.hdb_last_error().decode()
    rc = lib.hdb_cluster(None, p, 1, 8, 2, 0, 2, 2, 0.0, p, p, 16, p, 0)
    assert rc != 0 and "workspace" in lib.hdb_last_error().decode()
    rc = lib.hdb_cluster_matrix(None, p, None, 1, 8, 0, 2, 0, 0.0, -100000, p, p, 1 << 20, p, 0, None, None)
    assert rc != 0 and "core" in lib.hdb_last_error().decode()


def test_product_path_never_references_the_oracle():
    for folder in ("hdbscan_amd",):
        for dirpath, _, files in os.walk(os.path.join(ROOT, folder)):
            for f in files:
                if f.endswith((".py", ".hip", ".h", ".cpp")):
                    text = open(os.path.join(dirpath, f)).read()
                    assert "oracle" not in text.lower().replace("oracle (", ""), os.path.join(dirpath, f)


def test_noalloc_rejects_calls_above_its_scratch_before_any_device_work():
    lib = _native.lib()
    p = ctypes.c_void_p(1)
    assert lib.hdb_workspace_bytes(1024, 1024, 16, _native.F32, 5, 0) > _native.NOALLOC_BYTES
    rc = lib.hdb_cluster_noalloc(None, p, 1024, 1024, 16, _native.F32, 5, 5, 0.0, p, p, 0)
    assert rc != 0 and "scratch" in lib.hdb_last_error().decode()
    rc = lib.hdb_cluster_noalloc(None, None, 1, 8, 2, _native.F32, 2, 2, 0.0, p, p, 0)
    assert rc != 0 and "null" in lib.hdb_last_error().decode()
