"""The allocation-free row analysis (oc_analysis_noalloc) against
oc_analysis, through the C-ABI: every output bit-exact, on two calls in a
row (the second on a smaller input, so stale scratch contents from the
first call would show), plus the rejection of operands beyond the scratch
limit.  oc_analysis itself is pinned against the oracle by
test_gpu_stages.py::test_analysis."""
import ctypes as C

import numpy as np
import pytest

from helpers import log

pytestmark = pytest.mark.gpu


def _analysis(gpu, fn, dA, dB, m):
    outs = [gpu.DeviceBuffer(4 * max(m, 1)) for _ in range(5)]
    psum = gpu.DeviceBuffer(8 * (m + 1))
    sc = gpu.DeviceBuffer(8)
    for b in outs + [psum, sc]:
        b.zero()
    gpu._check(fn(C.byref(dA.m), C.byref(dB.m), *[o.ptr for o in outs], psum.ptr, sc.ptr,
                  sc.ptr + 4, None))
    gpu._check(gpu.lib.oc_device_synchronize())
    return ([o.to_array(np.uint32, m) for o in outs], psum.to_array(np.uint64, m + 1),
            sc.to_array(np.uint32, 2))


def _same(gpu, A, B, name):
    dA, dB = A.upload(), B.upload()
    want = _analysis(gpu, gpu.lib.oc_analysis, dA, dB, A.rows)
    got = _analysis(gpu, gpu.lib.oc_analysis_noalloc, dA, dB, A.rows)
    for k, (g, w) in enumerate(zip(got[0], want[0])):
        assert np.array_equal(g, w), f"{name}: per-row output {k} differs"
    assert np.array_equal(got[1], want[1]), f"{name}: product prefix differs"
    assert np.array_equal(got[2], want[2]), f"{name}: maxima differ"
    log(f"{name}: {A.rows} rows, total products {int(got[1][-1])}, maxima {got[2].tolist()}")


def test_analysis_noalloc_matches(gpu):
    """Two calls in a row, the second smaller; then a call with hub rows
    (the long-row list), then one with no rows."""
    log("case: analysis noalloc, rmat 2^16 then rmat 2^12")
    A = gpu.CSR.rmat(16, 16 << 16, seed=1)
    _same(gpu, A, A, "rmat16")
    A2 = gpu.CSR.rmat(12, 8 << 12, seed=2)
    _same(gpu, A2, A2, "rmat12 (smaller, second call)")
    rng = np.random.default_rng(4)
    n = 50000
    lens = rng.integers(0, 8, 300)
    lens[[3, 77, 150]] = [40000, 3000, 1025]
    rp = np.zeros(301, np.uint32)
    np.cumsum(lens, out=rp[1:])
    cols = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens])
    H = gpu.CSR(300, n, rp, cols.astype(np.uint32), rng.uniform(-1, 1, int(rp[-1])))
    Bh = gpu.CSR.erdos_renyi(n, 6.0, seed=5)
    _same(gpu, H, Bh, "hub rows")
    E = gpu.CSR(0, n, np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    _same(gpu, E, Bh, "no rows")


def test_analysis_noalloc_rejects_large(gpu):
    """B with 2^21 + 2 rows needs more than the 32 MiB of scratch (16 B per
    B row): rejected with OC_ERR_INVALID before any launch."""
    log("case: analysis noalloc beyond the scratch limit")
    nb = (1 << 21) + 2
    B = gpu.CSR(nb, 4, np.zeros(nb + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    A = gpu.CSR(2, nb, np.array([0, 1, 1], np.uint32), np.array([5], np.uint32), np.ones(1))
    dA, dB = A.upload(), B.upload()
    outs = [gpu.DeviceBuffer(8) for _ in range(5)]
    psum = gpu.DeviceBuffer(24)
    rc = gpu.lib.oc_analysis_noalloc(C.byref(dA.m), C.byref(dB.m), *[o.ptr for o in outs],
                                     psum.ptr, None, None, None)
    assert rc == gpu.OC_ERR_INVALID
    assert b"limit" in gpu.lib.oc_last_error()
