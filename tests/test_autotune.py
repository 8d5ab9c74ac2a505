"""The ragged autotuner's table handling on the CPU: flag encoding, the
table format, installing and clearing it through the C ABI."""

import ctypes
import json

import pytest

from hdbscan_amd import _native, autotune


def table(flags):
    return {"d": 16, "precision": "float32", "min_samples": 5,
            "buckets": [{"top": t, "chosen_flags": f} for t, f in zip(autotune.BUCKETS, flags)]}


def test_describe_decodes_every_tuning_field():
    f = (4 << 8) | (3 << 12) | (2 << 14) | _native.FLAG_GLOBAL_ARENA
    d = autotune.describe(f)
    assert d == {"vertices_per_thread": 4, "gram_row_sets": 4, "repack": 32, "tree_arena": "device",
                 "prim_points": "lds"}
    assert autotune.describe(0)["vertices_per_thread"] == "auto"


def test_flags_of_keeps_only_tuning_bits_and_orders_buckets():
    t = table([0, 1 << 8, 2 << 12 | _native.FLAG_MATRIX, 3 << 14, _native.FLAG_NO_POINT_LDS, 0, 4 << 8])
    t["buckets"].reverse()
    assert autotune.flags_of(t) == [0, 1 << 8, 2 << 12, 3 << 14, _native.FLAG_NO_POINT_LDS, 0, 4 << 8]
    with pytest.raises(ValueError):
        autotune.flags_of({"buckets": [{"top": 100, "chosen_flags": 0}]})


def test_save_load_apply_and_clear(tmp_path):
    path = str(tmp_path / "t.json")
    autotune.save(table([1 << 8] * 7), path)
    loaded = autotune.load(path)
    assert json.load(open(path)) == loaded
    autotune.apply(loaded)
    autotune.apply(None)
    with pytest.raises(ValueError, match="lacks"):
        autotune.save({"d": 1}, path)
        autotune.load(path)


def test_set_tuning_rejects_bad_counts():
    lib = _native.lib()
    assert lib.hdb_ragged_set_tuning(None, 8) != 0
    assert lib.hdb_ragged_set_tuning(None, 3) != 0
    assert lib.hdb_ragged_set_tuning(None, 0) == 0
    arr = (ctypes.c_int32 * 7)(*([0] * 7))
    assert lib.hdb_ragged_set_tuning(arr, 7) == 0
    assert lib.hdb_ragged_set_tuning(None, 0) == 0


def test_committed_tables_are_valid():
    import glob
    import os

    for path in glob.glob(os.path.join(autotune.HERE, "tuning", "*.json")):
        t = autotune.load(path)
        assert len(t["buckets"]) >= 1
        for entry in t["buckets"]:
            assert entry["chosen_ms"] <= entry["auto_ms"]
            assert any(m["flags"] == entry["chosen_flags"] for m in entry["measurements"])
