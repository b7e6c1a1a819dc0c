"""CPU oracle of the spy-plot segmentation (test infrastructure only).

numpy restatement of Segmentation.cuh:374-673 as the library computes it:
tile density of A's nonzeros (Segmentation.cuh:73-101), mean / deviation
and hot tiles (103-121, 430-479), 4-connected components labelled by the
smallest flat index of each (123-180; synchronous min-label sweeps whose
count is `cc_iterations`), the dominant label of each
grid row (182-243; ties to the smaller label), run-length encoding into row
bands, the merge of bands shorter than min_segment_rows (560-603), and the
per-segment sums and class rule (245-372).  The label propagation is
bounded by cc_max_iters sweeps and followed by up to 8 pointer-jumping
rounds, as in the reference (495-537), but synchronous, so it is
deterministic.

Pinned in tests/test_oracle_segment.py: tile counts against
numpy.histogram2d, components against scipy.ndimage.label, bands and
components of a hand-built two-block matrix, segment sums by brute force.
"""
from __future__ import annotations

import numpy as np

CLASSES = ("SPARSE_UNIFORM", "DENSE", "POWER_LAW", "BANDED")
EMPTY = 0xFFFFFFFF


def grid_shape(a_rows: int, b_cols: int, grid_max: int):
    g = max(8, grid_max)
    gh, gw = min(g, a_rows), min(g, b_cols)
    return gh, gw, -(-a_rows // gh), -(-b_cols // gw)


def density(A, gh, gw, tr, tc) -> np.ndarray:
    ro = A.row_offsets.astype(np.int64)
    rows = np.repeat(np.arange(A.rows, dtype=np.int64), np.diff(ro))
    t_r = np.minimum(rows // tr, gh - 1)
    t_c = np.minimum(A.col_ids.astype(np.int64) // tc, gw - 1)
    return np.bincount(t_r * gw + t_c, minlength=gh * gw).astype(np.uint64).reshape(gh, gw)


def hot_labels(d: np.ndarray, k: float):
    n = d.size
    s, sq = int(d.sum()), int((d.astype(np.uint64) ** 2).sum())
    mean = s / n
    var = max(sq / n - mean * mean, 0.0)
    std = var ** 0.5
    thresh = mean + k * std
    flat = d.ravel()
    hot = (flat.astype(np.float64) >= thresh) & (flat > 0)
    lab = np.where(hot, np.arange(n, dtype=np.int64), EMPTY).reshape(d.shape)
    return lab, mean, std, s, sq


def components(lab: np.ndarray, max_sweeps: int = 32):
    """At most max_sweeps synchronous 4-neighbour min-label sweeps, then up
    to 8 synchronous pointer-jumping rounds, each stage stopping once
    nothing changes; returns (labels, sweeps that changed a label)."""
    lab = lab.copy()
    hot = lab != EMPTY
    sweeps = 0
    for _ in range(max_sweeps):
        best = lab.copy()
        best[:, 1:] = np.minimum(best[:, 1:], lab[:, :-1])
        best[:, :-1] = np.minimum(best[:, :-1], lab[:, 1:])
        best[1:, :] = np.minimum(best[1:, :], lab[:-1, :])
        best[:-1, :] = np.minimum(best[:-1, :], lab[1:, :])
        best = np.where(hot, best, EMPTY)
        if np.array_equal(best, lab):
            break
        lab = best
        sweeps += 1
    flat = lab.ravel()
    for _ in range(8):
        nxt = np.where(flat == EMPTY, EMPTY, flat[np.where(flat == EMPTY, 0, flat)])
        if np.array_equal(nxt, flat):
            break
        flat = nxt
    return flat.reshape(lab.shape), sweeps


def dominant(lab: np.ndarray):
    out = np.full(lab.shape[0], EMPTY, dtype=np.int64)
    for g in range(lab.shape[0]):
        row = lab[g][lab[g] != EMPTY]
        if len(row):
            vals, cnt = np.unique(row, return_counts=True)
            out[g] = vals[np.argmax(cnt)]  # unique sorts: the first max is the smaller label
    return out


def bands(row_label, tile_rows, a_rows, min_rows):
    raw, cur, cur_g = [], row_label[0], 0
    for g in range(1, len(row_label)):
        if row_label[g] != cur:
            b0, b1 = cur_g * tile_rows, min(a_rows, g * tile_rows)
            if b1 > b0:
                raw.append([b0, b1, int(cur)])
            cur, cur_g = row_label[g], g
    if a_rows > cur_g * tile_rows:
        raw.append([cur_g * tile_rows, a_rows, int(cur)])
    merged = []
    for b in raw:
        if merged and b[1] - b[0] < max(1, min_rows):
            merged[-1][1] = b[1]
        else:
            merged.append(list(b))
    ncomp = len({b[2] for b in merged if b[2] != EMPTY})
    if not merged:
        merged = [[0, a_rows, EMPTY]]
    return merged, ncomp


def classify(seg_rows, A, an, b_cols, exp_threshold):
    """Per segment [start, end): the reference's sums and class rule."""
    ro = A.row_offsets.astype(np.int64)
    P = an["products"].astype(np.int64)
    out = []
    for s0, s1 in seg_rows:
        p = P[s0:s1]
        na = np.diff(ro[s0:s1 + 1])
        sp, sa, sq = int(p.sum()), int(na.sum()), int((p.astype(np.uint64) ** 2).sum())
        lo = int(an["left"][s0:s1].min()) if s1 > s0 else EMPTY
        hi = int(an["right"][s0:s1].max()) if s1 > s0 else 0
        rows = s1 - s0
        mean_p = sp / rows if rows else 0.0
        var_p = max(sq / rows - mean_p * mean_p, 0.0) if rows else 0.0
        cv = var_p ** 0.5 / mean_p if mean_p > 0 else 0.0
        er = sp / sa if sa else 0.0
        col_left = 0 if lo == EMPTY else lo
        span = (hi + 1 - col_left) & 0xFFFFFFFF
        ratio = span / b_cols
        row_density = mean_p / (b_cols * ratio) if ratio > 0 else 0.0
        if 0 < ratio < 0.05 and mean_p > 0 and er >= exp_threshold:
            cls = 3
        elif row_density >= 0.25:
            cls = 1
        elif cv >= 1.5:
            cls = 2
        else:
            cls = 0
        out.append({"row_start": s0, "row_end": s1, "row_count": rows, "col_left": col_left,
                    "col_right": hi, "class": CLASSES[cls], "sum_products": sp, "sum_a_nnz": sa,
                    "mean_row_len": mean_p, "var_row_len": var_p, "local_er": er,
                    "max_b_row_len": int(an["max_b"][s0:s1].max()) if rows else 0,
                    "max_a_row_len": int(na.max()) if rows else 0})
    return out


def segmentation(A, B, an, grid_max=256, k=2.0, min_rows=1024, exp_threshold=8, cc_max_iters=32):
    """The stats JSON `segmentation` block (times aside) for C = A*B."""
    gh, gw, tr, tc = grid_shape(A.rows, B.cols, grid_max)
    d = density(A, gh, gw, tr, tc)
    lab0, mean, std, _, _ = hot_labels(d, k)
    lab, sweeps = components(lab0, cc_max_iters)
    rl = dominant(lab)
    merged, ncomp = bands(rl, tr, A.rows, min_rows)
    segs = classify([(b[0], b[1]) for b in merged], A, an, B.cols, exp_threshold)
    return {"grid_h": gh, "grid_w": gw, "tile_rows": tr, "tile_cols": tc, "density_mean": mean,
            "density_stddev": std, "num_components": ncomp, "cc_iterations": sweeps,
            "segments": segs, "_density": d, "_labels": lab}
