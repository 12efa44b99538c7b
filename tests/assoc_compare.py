"""Association parity rule shared by the GPU tests and its own CPU check.

An association (row, target) set is compared with the oracle's
(fill_in_A_mat_gpu semantics, CvoGPU.cu:951-1099):

  * a pair whose oracle value lies within `near` relative of sp_thres, or whose
    distance to a gate threshold is within `near` relative, is "near": float
    rounding may decide it either way;
  * a row where neither side reached the cap K must have no pair outside the
    near ones in the symmetric difference;
  * a row where either side holds K pairs keeps the first K passing targets by
    index (CvoGPU.cu:994-996).  A near pair can shift which passing targets
    fill the last slots, so such a row is re-evaluated target by target (the
    oracle's pair values of every target up to the last one either side
    kept, every target when the GPU row is short of K): the GPU's targets must
    all pass or be near, every non-near passing target below its last kept
    one must be among them, and a GPU row short of K must hold every non-near
    passing target.  Without a near pair in the row that means an exact
    match.

Stored values of common pairs must agree within `val_rtol` relative, and each
GPU row must list its targets in ascending order.
"""
from __future__ import annotations

import numpy as np


def oracle_triplets(counts, cols, vals):
    """(rows, cols, vals) of the oracle's ELL rows."""
    K = cols.shape[1]
    mask = np.arange(K)[None, :] < counts[:, None]
    rows = np.nonzero(mask)[0]
    return rows, cols[mask].astype(np.int64), vals[mask]


def check_rows(want_rows, want_cols, want_vals, got_rows, got_cols, got_vals, n_src, n_tgt, K, near_of, sp,
               near=1e-6, val_rtol=4e-6):
    """Applies the rule.  near_of(i, j) -> (a, margin) arrays for the differing
    pairs (a < 0: the oracle rejects the pair).  Returns a dict of counts;
    raises AssertionError with the offending rows otherwise."""
    M = np.int64(n_tgt)
    wk = want_rows.astype(np.int64) * M + want_cols.astype(np.int64)
    gk = got_rows.astype(np.int64) * M + got_cols.astype(np.int64)
    only_w = np.setdiff1d(wk, gk, assume_unique=True)
    only_g = np.setdiff1d(gk, wk, assume_unique=True)
    w_cnt = np.bincount(want_rows, minlength=n_src)
    g_cnt = np.bincount(got_rows, minlength=n_src)
    bad = []
    if only_w.size or only_g.size:
        dk = np.concatenate([only_w, only_g])
        in_got = np.concatenate([np.zeros(only_w.size, bool), np.ones(only_g.size, bool)])
        di, dj = dk // M, dk % M
        a, margin = near_of(di, dj)
        is_near = ((a > 0) & (np.abs(a - sp) <= near * sp)) | (margin <= near)
        order = np.argsort(di, kind="stable")
        di, dj, in_got, is_near = di[order], dj[order], in_got[order], is_near[order]
        bounds = np.flatnonzero(np.diff(di)) + 1
        g_order = np.argsort(got_rows, kind="stable")
        g_start = np.concatenate([[0], np.cumsum(g_cnt)])
        w_order = np.argsort(want_rows, kind="stable")
        w_start = np.concatenate([[0], np.cumsum(w_cnt)])
        for seg in np.split(np.arange(di.size), bounds):
            i = int(di[seg[0]])
            far = ~is_near[seg]
            if not far.any():
                continue
            if not (w_cnt[i] >= K or g_cnt[i] >= K):
                bad.append((i, "uncapped row", dj[seg][far][:5].tolist()))
                continue
            G = np.sort(got_cols[g_order[g_start[i]:g_start[i + 1]]].astype(np.int64))
            W = want_cols[w_order[w_start[i]:w_start[i + 1]]].astype(np.int64)
            last = int(max(G.max(initial=-1), W.max(initial=-1)))
            J = n_tgt - 1 if g_cnt[i] < K else last
            js = np.arange(J + 1, dtype=np.int64)
            a_r, m_r = near_of(np.full(js.size, i, np.int64), js)
            near_r = ((a_r > 0) & (np.abs(a_r - sp) <= near * sp)) | (m_r <= near)
            pass_far = (a_r > sp) & ~near_r
            ok_g = pass_far | near_r
            if not ok_g[G].all():
                bad.append((i, "capped row keeps a target that does not pass", G[~ok_g[G]][:5].tolist()))
                continue
            gmax = int(G.max(initial=-1))
            must = js[pass_far & (js <= gmax)] if g_cnt[i] >= K else js[pass_far]
            missing = np.setdiff1d(must, G)
            if missing.size:
                bad.append((i, "capped row skips a passing target", missing[:5].tolist()))
    assert not bad, f"{len(bad)} rows break the association rule: {bad[:8]}"
    # stored values of the common pairs
    common, wi, gi = np.intersect1d(wk, gk, assume_unique=True, return_indices=True)
    if common.size:
        wv = want_vals[wi].astype(np.float64)
        gv = got_vals[gi].astype(np.float64)
        rel = float(np.max(np.abs(gv - wv) / wv))
    else:
        rel = 0.0
    assert rel <= val_rtol, rel
    return {"common": int(common.size), "diff": int(only_w.size + only_g.size), "rel": rel,
            "capped_rows": int(((w_cnt >= K) | (g_cnt >= K)).sum())}


def rows_ascending(row_ptr, col):
    """Every row of a CSR lists its targets in strictly ascending order."""
    if col.size < 2:
        return True
    d = np.diff(col.astype(np.int64))
    starts = np.asarray(row_ptr[1:-1], np.int64) - 1
    ok = d > 0
    ok[starts[(starts >= 0) & (starts < d.size)]] = True
    return bool(ok.all())


def compare_with_oracle(oracle, params, src, tgt, T, ell, gpu_assoc, K=None, near=1e-6, val_rtol=4e-6):
    """The GPU association (runtime.Association) against the oracle at the same
    pose and lengthscale."""
    K = K or params.nearest_neighbors_max
    counts, cols, vals, _ = oracle.assoc(params, src, tgt, T, ell, K)
    wr, wc, wv = oracle_triplets(counts, cols, vals)
    gr, gc, gv = gpu_assoc.pairs()

    def near_of(i, j):
        return oracle.pair_values(params, src, tgt, T, ell, i, j)

    out = check_rows(wr, wc, wv, np.asarray(gr), np.asarray(gc), np.asarray(gv), src.n, tgt.n, K, near_of,
                     params.sp_thres, near, val_rtol)
    assert rows_ascending(gpu_assoc.row_ptr, gpu_assoc.col)
    return out
