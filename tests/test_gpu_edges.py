"""Degenerate inputs through the whole product on the GPU, against the oracle:
no rows, no entries, a single entry, A entries that meet only empty B rows
(every row has zero products), a ragged mix of empty and full rows, and a
dimension mismatch, which is an error before anything is launched
(main.cu:48-51 checks the same condition)."""
import numpy as np
import pytest

from helpers import check_product, csr_from_entries as _csr, log

pytestmark = pytest.mark.gpu


def _cases(gpu):
    rng = np.random.default_rng(21)
    B54 = _csr(gpu, 5, 4, [(r, c, float(rng.uniform(-1, 1))) for r in range(5) for c in range(4)
                           if rng.random() < 0.6])
    # B rows 0, 2 and 4 empty
    Bhole = _csr(gpu, 6, 300, [(r, c, float(rng.uniform(-1, 1))) for r in (1, 3, 5)
                               for c in range(300) if rng.random() < 0.3])
    ragged = []
    for r in range(40):
        if r % 3 == 0:
            continue  # empty A row
        for c in range(6):
            if rng.random() < 0.5:
                ragged.append((r, c, float(rng.uniform(-1, 1))))
    return {
        "no_rows": (_csr(gpu, 0, 5, []), B54),
        "no_entries": (_csr(gpu, 6, 5, []), B54),
        "one_entry": (_csr(gpu, 1, 1, [(0, 0, 2.0)]), _csr(gpu, 1, 1, [(0, 0, 3.0)])),
        "only_empty_b_rows": (_csr(gpu, 8, 6, [(r, c, 1.0 + r) for r in range(8) for c in (0, 2, 4)]),
                              Bhole),
        "ragged": (_csr(gpu, 40, 6, ragged), Bhole),
    }


CASES = ["no_rows", "no_entries", "one_entry", "only_empty_b_rows", "ragged"]


@pytest.mark.parametrize("workflow", ["default", "precise", "estimation"])
@pytest.mark.parametrize("case", CASES)
def test_degenerate_products(gpu, case, workflow):
    log(f"case: degenerate product {case}, workflow {workflow}")
    A, B = _cases(gpu)[case]
    cfg = gpu.Config() if workflow == "default" else gpu.Config(use_ultrasparse_workflow=0)
    sp = gpu.SpGEMM(cfg)
    if workflow != "default":
        sp.set_workflow(0 if workflow == "precise" else 1)
    dA, dB = A.upload(), B.upload()
    C = sp.run(dA, dB).download()
    check_product(A, B, C, 1e-12, f"{case}/{workflow}")
    if case == "one_entry":
        assert C.nnz == 1 and C.values[0] == 6.0
    if case in ("no_rows", "no_entries", "only_empty_b_rows"):
        assert C.nnz == 0
    log(f"{case}/{workflow}: nnz(C) {C.nnz}, workflow {sp.get_stats()['decision']['workflow']}")


def test_dimension_mismatch_is_an_error(gpu):
    log("case: A.cols != B.rows")
    A = _csr(gpu, 3, 4, [(0, 1, 1.0), (2, 3, 2.0)])
    B = _csr(gpu, 5, 2, [(4, 1, 1.0)])
    sp = gpu.SpGEMM(gpu.Config())
    with pytest.raises(gpu.OceanError, match="Incompatible matrix dimensions"):
        sp.run(A.upload(), B.upload())
    # the handle stays usable
    C = sp.run(A.upload(), _csr(gpu, 4, 2, [(1, 0, 1.0), (3, 1, -1.0)]).upload()).download()
    assert C.nnz == 2
