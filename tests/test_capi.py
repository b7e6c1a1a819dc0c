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
