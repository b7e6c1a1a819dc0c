"""Oracle pinning: MurmurHash3 known answers and HyperLogLog invariants."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as ora
from golden.make_murmur_vectors import murmur3_x86_32

HERE = os.path.dirname(os.path.abspath(__file__))

# Published MurmurHash3_x86_32 known answers.
KATS = [(b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
        (b"\x00\x00\x00\x00", 0, 0x2362F9DE), (b"\xff\xff\xff\xff", 0, 0x76293B50),
        (b"\x21\x43\x65\x87", 0, 0xF55B516B), (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE),
        (b"aaaa", 0x9747B28C, 0x5A97808A), (b"abc", 0, 0xB3DD93FA), (b"a", 0x9747B28C, 0x7FA09EA6)]


@pytest.mark.parametrize("data,seed,expect", KATS)
def test_murmur3_kat(data, seed, expect):
    assert ora.murmur3(data, seed) == expect
    assert murmur3_x86_32(data, seed) == expect


def test_murmur3_col_vectors():
    vec = json.load(open(os.path.join(HERE, "golden", "murmur3_seed1234.json")))
    for c, h in zip(vec["cols"], vec["hashes"]):
        assert ora.murmur3_col(c) == h


def _csr_rows(rows):
    """Host CSR-like object from lists of sorted column ids."""
    class M:
        pass
    m = M()
    m.rows = len(rows)
    m.row_offsets = np.zeros(len(rows) + 1, dtype=np.uint32)
    m.row_offsets[1:] = np.cumsum([len(r) for r in rows])
    m.col_ids = np.array([c for r in rows for c in r], dtype=np.uint32)
    return m


@pytest.mark.parametrize("log_p", [5, 6])
def test_hll_registers_closed_form(log_p):
    p = 1 << log_p
    rows = [[], [7], list(range(0, 1000, 3)), [5, 5000, 70000]]
    regs = ora.hll_construct(log_p, _csr_rows(rows)).reshape(len(rows), p)
    assert not regs[0].any()
    for ri, r in enumerate(rows):
        want = np.zeros(p, dtype=np.uint8)
        for c in r:
            h = murmur3_x86_32(int(c).to_bytes(4, "little"), 1234)
            clz = 32 - h.bit_length()
            want[h & (p - 1)] = max(want[h & (p - 1)], min(clz + 1, 33 - log_p))
        assert np.array_equal(regs[ri], want)


@pytest.mark.parametrize("log_p", [5, 6])
def test_hll_estimate_cases(log_p):
    p = 1 << log_p
    lc = ora.hll_lc_table(log_p)
    # linear-counting table: int(P * logf(P / v)) in float arithmetic
    for v in range(1, p + 1):
        assert lc[v] == int(np.float32(p) * np.log(np.float32(p) / np.float32(v)).astype(np.float32))
    assert ora.hll_estimate(log_p, np.zeros(p, np.uint8)) == 0  # empty set
    # all registers equal to k: z = P 2^-k, E = alpha P 2^k
    alpha = {5: 0.697, 6: 0.709}[log_p]
    for k in (3, 6, 10):
        assert ora.hll_estimate(log_p, np.full(p, k, np.uint8)) == int(alpha * p * p / (p * 2.0 ** -k))


def test_hll_merge_invariants():
    rng = np.random.default_rng(0)
    rows = [sorted(set(rng.integers(0, 50000, size=rng.integers(1, 300)).tolist())) for _ in range(40)]
    B = _csr_rows(rows)
    hb = ora.hll_construct(6, B)
    # A row referencing B rows i and j: merged registers = byte max, which is
    # commutative, associative and idempotent; the estimate equals the estimate
    # of the sketch of the union of the column sets.
    A = _csr_rows([[0, 1], [1, 0], [0, 0, 1], [2, 3, 4]])
    est = ora.hll_merge(6, A, hb, 1.0, np.zeros(7, np.uint32))
    assert est[0] == est[1] == est[2]
    union = sorted(set(rows[2]) | set(rows[3]) | set(rows[4]))
    u = ora.hll_construct(6, _csr_rows([union]))
    assert est[3] == ora.hll_estimate(6, u)


def test_hll_accuracy_statistical():
    # relative standard error 1.04/sqrt(P) on large random sets
    rng = np.random.default_rng(1)
    errs = []
    for _ in range(64):
        n = 20000
        cols = sorted(set(rng.integers(0, 1 << 30, size=n).tolist()))
        regs = ora.hll_construct(6, _csr_rows([cols]))
        errs.append(ora.hll_estimate(6, regs) / len(cols) - 1.0)
    assert abs(np.mean(errs)) < 3 * 1.04 / math.sqrt(64) / math.sqrt(64) + 0.02
    assert np.std(errs) < 2 * 1.04 / math.sqrt(64)


def test_hll_expansion_rounding():
    sizes = np.array([0, 331, 673, 1327, 2719, 5449, 16253], np.uint32)  # reference A100 table
    B = _csr_rows([list(range(0, 4000, 2))])
    A = _csr_rows([[], [0]])
    hb = ora.hll_construct(5, B)
    est = ora.hll_merge(5, A, hb, 2.0, sizes)
    assert est[0] == 0
    raw = ora.hll_estimate(5, hb[:32])
    t = int(raw * 2.0)
    want = next((int(s) for s in sizes[:6] if t <= s), t)
    assert est[1] == want


def test_sample_rows_lcg():
    # x = 1664525 * (112233 + i * 2654435761) + 1013904223 mod 2^32; row = x mod rows
    for i in (0, 1, 7, 10799):
        x = (1664525 * (112233 + i * 2654435761) + 1013904223) % (1 << 32)
        assert ora.sample_row(i, 1 << 20) == x % (1 << 20)
        assert ora.sample_row(i, 65536) == x % 65536
