"""Oracle pinning: binning rules and workflow decisions on hand-derived
cases, with the reference's own A100 tables (include/Common.h:115-269)."""
import math

import numpy as np

from oracle import oracle as ora

# A100 / 128 KB tables of the reference
HASH = [503, 1013, 2029, 4079, 8179, 16369, 16369]
DENSE = [0, 32736, 65504, 131040, 262112, 524256, 0]
HASH_NUM = [0, 331, 673, 1327, 2719, 5449, 16253]
DENSE_NUM = [236, 451, 906, 1816, 3636, 7277, 0xFFFFFFFF]
EMPTY_L, EMPTY_R = 0xFFFFFFFF, 0


def sym(products, left, right, comp=0, c=0.0):
    return list(ora.bin_symbolic(products, left, right, HASH, DENSE, 5, comp, 1.25, c))


def test_symbolic_bins():
    got = sym([0, 402, 404, 20000, 20000, 8000],
              [EMPTY_L, 0, 0, 0, 0, 0],
              [EMPTY_R, 99999, 19999, 599999, 499999, 999999])
    # empty row -> hash 0; 402*1.25=502 -> hash 0 (dense 3 > 0); 404 -> 505,
    # hash 1, dense 1 <= 1 -> 1+7; 25000 > 16369 -> hash 6, span > 524256 -> 6;
    # span 500000 -> dense 5 <= 6 -> min(6,5)+7 = 12; 8000 -> 10000 -> hash 5
    assert got == [0, 0, 8, 6, 12, 5]
    # with compaction 4.0: 8000/4 = 2000 -> 2500 -> hash 3
    assert sym([8000], [0], [999999], comp=1, c=4.0) == [3]


def test_symbolic_bins_gfx950_rule():
    # max_dense_bin 6 (a dense bin-6 kernel exists): hash bin 6 -> dense 13
    got = ora.bin_symbolic([20000], [0], [499999], HASH, DENSE[:6] + [1 << 20], 6, 0, 1.25, 0.0)
    assert list(got) == [13]


def test_numeric_bins():
    inp = [0, 200, 3000, 5000]
    left = [EMPTY_L, 0, 0, 0]
    right = [EMPTY_R, 4999, 4999, 999999]
    bins, _ = ora.bin_numeric(inp, left, right, HASH_NUM, DENSE_NUM, 1, 1.5)
    assert list(bins) == [7, 1, 12, 13]


def test_numeric_hybrid_cap_in_row_order():
    bins, _ = ora.bin_numeric([5000, 5000, 5000], [0] * 3, [999999] * 3, HASH_NUM, DENSE_NUM,
                              1, 1.5, use_largest=1, hybrid_max=1)
    assert list(bins) == [6, 13, 13]


def test_numeric_est_modifies_dense0():
    bins, est = ora.bin_numeric([100, 100], [0, 0], [199, 999999], HASH_NUM, DENSE_NUM,
                                0, 0.0, modify_dense_est=1)
    assert list(bins) == [7, 1] and list(est) == [236, 100]


def test_decide_analysis():
    c = ora.decide_cfg()
    d = ora.decide_analysis(c, 65536, 65536, 326656, 65536, 65536, 1629192, 25)
    assert (d.ana1_type, d.spark_kernel_size, d.spark_need_outlier) == (0, 32, 0)
    d = ora.decide_analysis(c, 65536, 65536, 326656, 65536, 65536, 1629192, 40)
    assert (d.spark_need_outlier, d.spark_outlier_kernel_size) == (1, 64)
    d = ora.decide_analysis(c, 10, 5000, 100, 5000, 20, 1000, 50)
    assert d.ana1_type == 2  # short-fat x tall-skinny
    d = ora.decide_analysis(c, 1000, 1000, 5000, 1000, 1000, 100000, 500)
    assert (d.ana1_type, d.hll_log2_precision, d.hll_expansion_factor) == (1, 5, 2.0)
    d = ora.decide_analysis(c, 1000, 1000, 5000, 1000, 1000, 300000, 500)
    assert (d.hll_log2_precision, d.hll_expansion_factor) == (6, 1.5)
    assert ora.num_samples(c, 1 << 20) == 10800 and ora.num_samples(c, 1000) == 648
    assert ora.num_samples(c, 100000) == 3000


def test_decide_sampling():
    c = ora.decide_cfg()
    d = ora.Decision()
    ora.decide_sampling(c, 50, 1000, [100, 200], [10, 20], d)
    assert d.compaction_1 == 10.0 and d.compaction_2 == 10.0 and d.ana2_type == 1
    # total products arrive truncated to 32 bits: (2^32 + 100) / 50 = 2 < 8
    d = ora.Decision()
    ora.decide_sampling(c, 50, (1 << 32) + 100, [100, 200], [10, 20], d)
    assert d.ana2_type == 0
    # an empty sampled row gives 0/0 = NaN in compaction_2, which fails no
    # threshold and keeps the dense bitmap query
    d = ora.Decision()
    ora.decide_sampling(c, 50, 1000, [100, 0], [10, 0], d)
    assert math.isnan(d.compaction_2) and d.ana2_type == 1 and d.dense_query_bitmap == 1
    # precise path: variance and safe compaction
    d = ora.Decision()
    ora.decide_sampling(c, 50, 1000, [10, 20, 30], [10, 10, 10], d)
    assert d.ana2_type == 0 and d.compaction_avg == 2.0 and d.compaction_var == 1.0
    assert abs(d.safe_compaction - (2.0 - 2.53) / 2) < 1e-15 and d.use_estimated_symbolic == 0
    d = ora.Decision()
    ora.decide_sampling(c, 50, 1000, [40, 40, 40, 41], [10, 10, 10, 10], d)
    assert d.ana2_type == 0 and d.safe_compaction >= 1.5 and d.use_estimated_symbolic == 1
