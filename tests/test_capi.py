"""The C-ABI library loads and exports every function declared in
include/ocean_capi.h; error reporting works without a GPU."""
import ctypes
import os
import re

import research_amd as ra
from research_amd import ocean

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ocean_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oc_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(ocean.library_path)
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_reporting_without_gpu():
    rc = ocean.lib.oc_csr_load(b"/nonexistent/x.csr", -1, ctypes.byref(ocean._Csr()))
    assert rc == ocean.OC_ERR_IO
    assert b"could not open" in ocean.lib.oc_last_error()
    rc = ocean.lib.oc_config_from_json(ctypes.byref(ocean._Config()), b"{")
    assert rc == ocean.OC_ERR_INVALID


def test_tables_gfx950():
    t = ocean.tables()
    # 160 KB of LDS per CU: per-bin LDS budgets and the table sizes they give
    assert t["lds_bytes"] == [5120, 5120, 10240, 20480, 40960, 81920, 163840]
    assert t["hash_sym"][5] == (81920 - 256) // 4 // 8 * 8
    assert t["dense_sym"][6] == (163840 - 512) * 8 and t["dense_sym"][0] == 0
    assert t["hash_num"][0] == 0 and t["dense_num"][6] == 0xFFFFFFFF
    assert all(t["hash_num"][b] < t["hash_num"][b + 1] for b in range(1, 6))


def test_config_range_checks():
    """Expansion coefficients below 1 (a bin smaller than its row's output),
    a zero fat ratio (division), unknown HyperLogLog precisions and kernel
    knobs outside their ranges are rejected by the check, by the JSON reader
    and by handle creation."""
    def cfg(**kw):
        c = ocean._Config()
        ocean.lib.oc_config_default(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c
    assert ocean.lib.oc_config_check(ctypes.byref(cfg())) == ocean.OC_OK
    for kw in (dict(symbolic_expansion_coe=0.9), dict(numeric_expansion_coe=0.5),
               dict(fat_ratio=0), dict(hll_precision_high=7), dict(hll_sample_min_rows=0),
               dict(tile_block=300), dict(tile_log2_cols=20), dict(band_log2_cols=7)):
        assert ocean.lib.oc_config_check(ctypes.byref(cfg(**kw))) == ocean.OC_ERR_INVALID, kw
    for text in (b'{"tile_block": 300}', b'{"band_log2_cols": 7}'):
        assert ocean.lib.oc_config_from_json(ctypes.byref(cfg()), text) == ocean.OC_ERR_INVALID
    assert not ocean.lib.oc_spgemm_create(ctypes.byref(cfg(symbolic_expansion_coe=0.75)))
    assert b"symbolic_expansion_coe" in ocean.lib.oc_last_error()


def test_bin_shapes_config():
    """Per-bin shapes: JSON arrays of 7 integers, range checks, and tables
    that must fit each bin's LDS (checked when the handle is created)."""
    c = ocean.Config.from_json('{"bin_concurrency": [0, 0, 8, 0, 0, 0, 0], '
                               '"bin_block": [0, 128, 0, 0, 0, 1024, 0], "tile_block": 512}')
    assert c.bin_concurrency == [0, 0, 8, 0, 0, 0, 0]
    assert c.bin_block == [0, 128, 0, 0, 0, 1024, 0] and c.tile_block == 512
    base = ocean._Config()
    ocean.lib.oc_config_default(ctypes.byref(base))
    for text in (b'{"bin_block": [0, 0, 0, 300, 0, 0, 0]}', b'{"bin_concurrency": [65, 0, 0, 0, 0, 0, 0]}',
                 b'{"bin_block": [0, 0, 0]}'):
        assert ocean.lib.oc_config_from_json(ctypes.byref(base), text) == ocean.OC_ERR_INVALID, text
    # nested structures other than flat number arrays are still skipped
    fresh = ocean._Config()
    ocean.lib.oc_config_default(ctypes.byref(fresh))
    assert ocean.lib.oc_config_from_json(ctypes.byref(fresh), b'{"x": [1, [2]], "y": {"z": 1}}') == \
        ocean.OC_OK
    # 64 workgroups per CU leave 2.5 KB each: too little for a 1024-thread
    # numeric hash layout
    bad = ocean.Config(bin_concurrency=[0, 0, 0, 0, 0, 0, 64], bin_block=[0, 0, 0, 0, 0, 0, 1024])
    assert ocean.lib.oc_config_check(ctypes.byref(bad._c)) == ocean.OC_OK
    assert not ocean.lib.oc_spgemm_create(ctypes.byref(bad._c))
    assert b"bin 6" in ocean.lib.oc_last_error()
