"""MatrixMarket input: oc_mtx_load and the convert CLI against the
reference's own converter (utils/convert.cpp, built into oracle/_ref when the
reference tree is present) and against scipy's MatrixMarket reader."""
import os
import subprocess
import zlib

import numpy as np
import pytest
import scipy.io
import scipy.sparse

import research_amd as ra
from research_amd import ocean

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_CONVERT = os.path.join(ROOT, "oracle", "_ref", "convert")
CONVERT = os.path.join(ROOT, "research_amd", "convert")


def _write_mtx(path, rows, cols, entries, field="real", sym="general", blanks=True):
    """entries: (row, col, value) 1-based, written in the given order, with
    (blanks) an empty line, a comment and a whitespace-only line halfway."""
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {sym}\n")
        f.write("% a comment line\n")
        f.write(f"{rows} {cols} {len(entries)}\n")
        for i, (r, c, v) in enumerate(entries):
            if blanks and i == len(entries) // 2:
                f.write("\n% another comment\n   \n")
            if field == "pattern":
                f.write(f"{r} {c}\n")
            elif field == "integer":
                f.write(f"{r} {c} {int(v)}\n")
            elif field == "complex":
                f.write(f"{r} {c} {v!r} {(-v)!r}\n")
            else:
                f.write(f"  {r}\t{c} {v!r}\n")


def _random_entries(rng, rows, cols, n, sym=False, integer=False):
    seen, out = set(), []
    while len(out) < n:
        r, c = int(rng.integers(1, rows + 1)), int(rng.integers(1, cols + 1))
        if sym and r < c:
            r, c = c, r
        if (r, c) in seen:
            continue
        seen.add((r, c))
        v = float(rng.integers(-9, 10)) if integer else float(rng.uniform(-2, 2))
        out.append((r, c, v))
    return out


CASES = [
    ("real", "general", "double"),
    ("real", "general", "float"),
    ("real", "symmetric", "double"),
    ("integer", "general", "double"),
    ("pattern", "general", "float"),
    ("pattern", "symmetric", "double"),
    ("complex", "Hermitian", "double"),
]


@pytest.mark.skipif(not os.path.exists(REF_CONVERT), reason="reference converter not built here")
@pytest.mark.parametrize("field,sym,dtype", CASES)
def test_convert_matches_reference_converter(tmp_path, field, sym, dtype):
    """Same bytes as the reference's converter (the State padding, which the
    reference leaves uninitialised, aside) and the same stdout lines, on
    shuffled entries without duplicates."""
    rng = np.random.default_rng(zlib.crc32(f"{field} {sym} {dtype}".encode()))
    rows, cols = (57, 57) if sym != "general" else (61, 45)
    ent = _random_entries(rng, rows, cols, 300, sym != "general", field == "integer")
    rng.shuffle(ent)
    mtx = str(tmp_path / "a.mtx")
    _write_mtx(mtx, rows, cols, ent, field, sym)
    ref, mine = str(tmp_path / "ref.csr"), str(tmp_path / "mine.csr")
    pr = subprocess.run([REF_CONVERT, mtx, ref, dtype], capture_output=True, text=True, check=True)
    pm = subprocess.run([CONVERT, mtx, mine, dtype], capture_output=True, text=True, check=True)
    r, m = open(ref, "rb").read(), open(mine, "rb").read()
    assert len(r) == len(m)
    pad = (89, 96) if dtype == "double" else (85, 88)
    assert r[:pad[0]] == m[:pad[0]] and r[pad[1]:] == m[pad[1]:]
    assert m[pad[0]:pad[1]] == b"\0" * (pad[1] - pad[0])
    assert pr.stdout.replace(ref, mine).splitlines() == pm.stdout.splitlines()


@pytest.mark.parametrize("sym", ["general", "symmetric"])
@pytest.mark.parametrize("vtype", [ocean.OC_F64, ocean.OC_F32])
def test_load_mtx_against_scipy(tmp_path, sym, vtype):
    """Duplicates summed (merge_duplicates) equals scipy's reading of the
    same file converted to CSR (which sums duplicates and sorts columns)."""
    rng = np.random.default_rng(7 + vtype)
    n = 80
    ent = _random_entries(rng, n, n, 500, sym == "symmetric")
    ent += [ent[i] for i in rng.choice(len(ent), 60, replace=False)]  # duplicates
    rng.shuffle(ent)
    mtx = str(tmp_path / "d.mtx")
    _write_mtx(mtx, n, n, ent, "real", sym, blanks=False)  # scipy rejects them
    want = scipy.sparse.csr_matrix(scipy.io.mmread(mtx))
    want.sum_duplicates()
    want.sort_indices()
    got = ra.CSR.load_mtx(mtx, vtype, merge_duplicates=True)
    assert np.array_equal(got.row_offsets.astype(np.int64), want.indptr)
    assert np.array_equal(got.col_ids.astype(np.int64), want.indices)
    dt = np.float32 if vtype == ocean.OC_F32 else np.float64
    assert got.values.dtype == dt
    if vtype == ocean.OC_F64:
        assert np.allclose(got.values, want.data, rtol=1e-15, atol=1e-15)
    else:  # each value rounded to f32 once, then summed in f64 and rounded
        assert np.allclose(got.values, want.data.astype(np.float32), rtol=1e-6, atol=1e-6)
    kept = ra.CSR.load_mtx(mtx, vtype)
    assert kept.nnz == (len(ent) if sym == "general" else
                        sum(1 if r == c else 2 for r, c, _ in ent))


def test_duplicates_keep_file_order(tmp_path):
    mtx = str(tmp_path / "k.mtx")
    _write_mtx(mtx, 3, 4, [(2, 3, 5.0), (1, 4, 1.0), (2, 3, -7.0), (2, 1, 2.0), (2, 3, 0.5)])
    A = ra.CSR.load_mtx(mtx)
    assert list(A.row_offsets) == [0, 1, 5, 5]
    assert list(A.col_ids) == [3, 0, 2, 2, 2]
    assert list(A.values) == [1.0, 2.0, 5.0, -7.0, 0.5]
    B = ra.CSR.load_mtx(mtx, merge_duplicates=True)
    assert list(B.row_offsets) == [0, 1, 3, 3] and list(B.col_ids) == [3, 0, 2]
    assert list(B.values) == [1.0, 2.0, -1.5]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate form"),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.0\n", "symmetric, general"),
    ("%%MatrixMarket matrix coordinate text general\n2 2 1\n1 1 1.0\n", "data type"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "Row index out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 1.0\n", "Column index out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1.0\n", "Row index out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 x\n", "Failed to read data at line 2"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n2 2 1.0\n", "more entries"),
    ("%%MatrixMarket matrix coordinate real general\n% only comments\n", "header"),
])
def test_mtx_errors(tmp_path, text, msg):
    p = str(tmp_path / "e.mtx")
    open(p, "w").write(text)
    with pytest.raises(ocean.OceanError, match=msg):
        ra.CSR.load_mtx(p)


def test_convert_cli_arguments(tmp_path):
    mtx = str(tmp_path / "c.mtx")
    _write_mtx(mtx, 3, 3, [(1, 1, 1.0), (3, 2, 2.0), (1, 1, 4.0)])
    p = subprocess.run([CONVERT, mtx], capture_output=True, text=True)
    assert p.returncode == 1 and p.stdout.startswith("Usage:")
    p = subprocess.run([CONVERT, mtx, str(tmp_path / "o.csr"), "half"], capture_output=True, text=True)
    assert p.returncode == 1 and "Invalid data type: half" in p.stderr
    p = subprocess.run([CONVERT, str(tmp_path / "missing.mtx"), str(tmp_path / "o.csr")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and p.stderr.startswith("Error: could not open")
    out = str(tmp_path / "m.csr")
    subprocess.run([CONVERT, mtx, out, "float", "--merge-duplicates"], check=True, capture_output=True)
    M = ra.CSR.load(out, ocean.OC_F32)
    assert list(M.row_offsets) == [0, 1, 1, 2] and list(M.col_ids) == [0, 1]
    assert list(M.values) == [5.0, 2.0]
