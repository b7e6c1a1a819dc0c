"""Host-side surface: .csr files, compare(), config JSON, generators."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import research_amd as ra
from research_amd import ocean

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_CONVERT = os.path.join(ROOT, "oracle", "_ref", "convert")


def _mat(seed=0, vtype=np.float64):
    rng = np.random.default_rng(seed)
    m, n = 37, 53
    rows = [np.sort(rng.choice(n, size=rng.integers(0, 9), replace=False)) for _ in range(m)]
    rp = np.zeros(m + 1, np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.uint32)
    return ra.CSR(m, n, rp, ci, rng.uniform(-1, 1, len(ci)).astype(vtype))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr_roundtrip(tmp_path, dt):
    A = _mat(1, dt)
    p = str(tmp_path / "a.csr")
    A.store(p)
    raw = open(p, "rb").read()
    assert raw[:9] == b"Hi\x01Compsd" and len(raw) == 80 + (16 if dt == np.float64 else 8) + A.nnz * (
        np.dtype(dt).itemsize + 4) + (A.rows + 1) * 4
    hdr = np.frombuffer(raw[16:80], np.uint64)
    assert list(hdr) == [np.dtype(dt).itemsize, 0, 4, 0, 4, A.rows, A.cols, A.nnz]
    B = ra.CSR.load(p)
    assert np.array_equal(B.row_offsets, A.row_offsets) and np.array_equal(B.col_ids, A.col_ids)
    assert np.array_equal(B.values, A.values) and B.values.dtype == dt
    with pytest.raises(ocean.OceanError):  # typesize mismatch is an error (src/CSR.cpp:136-139)
        ra.CSR.load(p, 1 if dt == np.float64 else 0)


def test_csr_wide_offsets_header(tmp_path):
    # offsets are written as u64 only when nnz >= 2^32; u32 files otherwise
    A = _mat(2)
    A.row_offsets = A.row_offsets.astype(np.uint64)
    p = str(tmp_path / "w.csr")
    A.store(p)
    assert np.frombuffer(open(p, "rb").read()[48:56], np.uint64)[0] == 4
    assert np.array_equal(ra.CSR.load(p).row_offsets, A.row_offsets.astype(np.uint32))


@pytest.mark.skipif(not os.path.exists(REF_CONVERT), reason="reference converter not built here")
def test_against_reference_converter(tmp_path):
    """The reference's own .mtx -> .csr converter (built from
    utils/convert.cpp into oracle/_ref) and this library read and write the
    same bytes (the State padding, uninitialised in the reference, aside)."""
    A = _mat(3)
    mtx = tmp_path / "a.mtx"
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.rows} {A.cols} {A.nnz}\n")
        for r in range(A.rows):
            for j in range(A.row_offsets[r], A.row_offsets[r + 1]):
                f.write(f"{r + 1} {A.col_ids[j] + 1} {float(A.values[j])!r}\n")
    ref = str(tmp_path / "ref.csr")
    subprocess.run([REF_CONVERT, str(mtx), ref, "double"], check=True, capture_output=True)
    B = ra.CSR.load(ref, 0)
    assert np.array_equal(B.row_offsets, A.row_offsets) and np.array_equal(B.col_ids, A.col_ids)
    assert np.array_equal(B.values, A.values)
    mine = str(tmp_path / "mine.csr")
    A.store(mine)
    r, m = open(ref, "rb").read(), open(mine, "rb").read()
    assert len(r) == len(m)
    assert r[:89] == m[:89] and r[96:] == m[96:]  # bytes 89..95: State padding


def test_compare_semantics():
    A = _mat(4)
    ok, res = A.compare(A)
    assert ok and res["num_errors"] == 0
    B = ra.CSR(A.rows, A.cols, A.row_offsets.copy(), A.col_ids.copy(), A.values.copy())
    B.values[3] *= 1 + 5e-6
    ok, res = A.compare(B)
    assert ok and res["num_errors"] == 1  # counted above 1e-6, tolerated below 1e-4
    B.values[3] *= 1.01
    ok, res = A.compare(B)
    assert not ok and res["kind"] == 4
    C = ra.CSR(A.rows, A.cols, A.row_offsets.copy(), A.col_ids.copy(), A.values.copy())
    C.col_ids[0] += 1
    ok, res = A.compare(C)
    assert not ok and res["kind"] == 3


@pytest.mark.parametrize("name", ["analysis", "baseline", "bench", "bench_detail", "bench_seg"])
def test_reference_config_files(name):
    path = os.path.join(HERE, "golden", "config", name + ".json")
    c = ra.Config.load(path)
    d = json.load(open(path))
    key_to_field = {"track_stage_time": "track_stage_times"}
    for k, v in d.items():
        got = getattr(c, key_to_field.get(k, k))
        assert (got == v) if not isinstance(v, bool) else (bool(got) == v), k


def test_config_defaults_and_unknown_keys():
    c = ra.Config()
    assert (c.warmup_iters, c.bench_iters, c.hll_numeric_malloc_expansion_coe) == (10, 10, 1.3)
    assert c.use_estimation_workflow == 1 and c.use_segmented_workflow == 0
    c = ra.Config.from_json('{"bench_iters": 3, "no_such_key": [1, {"a": 2}], "print_info": true}')
    assert c.bench_iters == 3 and c.print_info == 1
    with pytest.raises(ocean.OceanError):
        ra.Config.from_json('{"bench_iters": "three"}')
    with pytest.raises(ocean.OceanError):
        ra.Config.from_json('[1, 2]')


def test_generators():
    L = ra.CSR.laplacian2d(256, 256, seed=0)
    assert L.nnz == 326656 and L.rows == 65536
    assert np.all(np.diff(L.row_offsets.astype(np.int64)) >= 3)
    assert np.all(L.values >= -1) and np.all(L.values < 1)
    R1, R2 = ra.CSR.rmat(12, 16 << 12, seed=1), ra.CSR.rmat(12, 16 << 12, seed=1)
    assert np.array_equal(R1.col_ids, R2.col_ids) and np.array_equal(R1.values, R2.values)
    for r in range(R1.rows):  # duplicates merged, columns ascending
        seg = R1.col_ids[R1.row_offsets[r]:R1.row_offsets[r + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
    assert R1.nnz < 16 << 12
    X = ra.CSR.rect_band(1 << 12, 1 << 15, 8.0, 0.3, 4096, 8, seed=2)
    assert X.values.dtype == np.float32 and abs(X.nnz / X.rows - 8.0) < 0.5
    XT = X.transpose()
    assert np.array_equal((X.to_scipy().T.tocsr() != XT.to_scipy()).nnz, 0)
    B = ra.CSR.banded(100, 5)
    assert B.nnz == 100 * 5 - 2 * (1 + 2)
