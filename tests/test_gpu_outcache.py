"""Product outputs come from a per-device cache of freed buffers
(out_alloc / out_free in host_csr.cpp): a freed output is reused best fit
for the next product, so buffers arrive holding an earlier product's
entries.  A series of products of different sizes, each freed before the
next, must still equal the oracle (structure bit-exact, values within
1e-12 condition-normalised), on one handle and on several."""
import numpy as np
import pytest

from helpers import check_product, log

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _series(gpu):
    # sizes that go up, down and up again, so later outputs land in larger
    # and smaller cached buffers
    return [gpu.CSR.rmat(13, 16 << 13, seed=21), gpu.CSR.erdos_renyi(3000, 6.0, seed=22),
            gpu.CSR.rmat(14, 24 << 14, seed=23), gpu.CSR.banded(5000, 17, seed=24),
            gpu.CSR.rmat(12, 8 << 12, seed=25), gpu.CSR.rmat(14, 24 << 14, seed=26)]


def test_output_cache_reuse_one_handle(gpu):
    log("case: output cache reuse, one handle, six products")
    sp = gpu.SpGEMM(gpu.Config())
    for i, A in enumerate(_series(gpu)):
        dA = A.upload()
        dC = sp.run(dA, dA)
        C = dC.download()
        dC.free()
        worst, _ = check_product(A, A, C, TOL64, f"product {i}")
        log(f"product {i}: n {A.rows}, nnz(C) {C.nnz}, err {worst:.2e}")


def test_output_cache_outputs_alive_together(gpu):
    """Outputs kept alive at once never share a buffer: two handles, the
    first product's output still held while the second runs."""
    log("case: output cache, outputs held together")
    mats = _series(gpu)
    sp1, sp2 = gpu.SpGEMM(gpu.Config()), gpu.SpGEMM(gpu.Config())
    held = []
    for i, A in enumerate(mats):
        dA = A.upload()
        held.append((A, (sp1 if i % 2 else sp2).run(dA, dA)))
    for i, (A, dC) in enumerate(held):
        worst, _ = check_product(A, A, dC.download(), TOL64, f"held product {i}")
        dC.free()
        log(f"held product {i}: err {worst:.2e}")
