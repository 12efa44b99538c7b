"""The oracle's adaptive-lengthscale CVO (AdaptiveCvoGPU::align,
src/cvo/AdaptiveCvoGPU.cu:1077-1248) against independent restatements:

  * its association (:263-408) against a numpy FP64 brute force: fixed
    lengthscale (the scene lies ~60 m out, where a range-scaled one differs
    by 13%), colour gate always on, the class factor without a gate, the
    first K passing targets by index (K = 6 sits below most rows' counts);
  * the lengthscale step: dl (:548-890) is the derivative in ell of
    <x,x> + <y,y> - 2<x,y> over the stored pairs divided by
    nnz(Axx) + nnz(Ayy) - 2 nnz(A) -- checked against a central finite
    difference of the three FP64 inner products -- and ell_1 = ell_0 -
    dl_step dl, or 0.7 ell_max past ell_max (:1198-1203), read from the trace;
  * the step rule (:1039-1074) against numpy.roots: smallest positive real
    root of 4E s^3 + 3D s^2 + 2C s + B, min_step when none, at most 0.8;
  * both precisions register the planes/spheres pair.
"""
import math

import numpy as np
import pytest

from conftest import params_from_yaml, rot_dist_rad, trans_err
from knn_cases import knn_pair, knn_params
from rkhs_ba_amd import synth


def moved(T, y):
    R, t = T[:3, :3].astype(np.float64), T[:3, 3].astype(np.float64)
    return (y.astype(np.float64) - t) @ R


def brute_row(p, xi, fxi, sxi, y, fy, sy, ell, sem):
    """a, d2 and the gate margins of one row against every target, FP64."""
    sig2 = float(p.sigma) ** 2
    thr = -2.0 * ell * ell * math.log(p.sp_thres / sig2)
    c2, cs2 = float(p.c_ell) ** 2, float(p.c_sigma) ** 2
    thr_c = -2.0 * c2 * math.log(p.sp_thres / cs2)
    s2, ss2 = float(p.s_ell) ** 2, float(p.s_sigma) ** 2
    d2 = ((y - xi) ** 2).sum(1)
    dc = ((fy.astype(np.float64) - fxi) ** 2).sum(1)
    a = sig2 * np.exp(-d2 / (2 * ell * ell)) * cs2 * np.exp(-dc / (2 * c2))
    if sem:
        ds = ((sy.astype(np.float64) - sxi) ** 2).sum(1)
        a = a * ss2 * np.exp(-ds / (2 * s2))
    ok = (d2 < thr) & (dc < thr_c) & (a > p.sp_thres)
    # relative distance to the closest gate, for pairs decided by rounding
    margin = np.minimum(np.minimum(np.abs(d2 - thr) / thr, np.abs(dc - thr_c) / thr_c),
                        np.abs(a - p.sp_thres) / p.sp_thres)
    return a, d2, ok, margin


def brute_pairs(p, x, fx, sx, y, fy, sy, ell, K, sem):
    """Rows of (j, a, d2) in ascending j, FP64."""
    rows = []
    for i in range(len(x)):
        a, d2, ok, _ = brute_row(p, x[i], fx[i], None if sx is None else sx[i], y, fy, sy, ell, sem)
        idx = np.nonzero(ok)[0][:K]
        rows.append([(int(j), a[j], d2[j]) for j in idx])
    return rows


@pytest.fixture(scope="module")
def far():
    p = knn_params(params_from_yaml("planes_spheres_5k.yaml"))
    p.is_using_geometric_type = 1  # ignored by this association
    src, tgt, T = knn_pair(0, n_src=300, n_tgt=1500, n_dup=0)
    return p, src, tgt, T


@pytest.mark.parametrize("K,ell", [(6, 0.5), (512, 0.3)])
def test_adaptive_association_matches_brute_force(oracle, far, K, ell):
    p, src, tgt, T = far
    # the oracle maps the target in float, as the reference does: 60 m out
    # a coordinate carries ~4e-6 m of rounding, so values agree to 1e-4 and a
    # pair within 1e-4 of a gate may fall either way
    counts, cols, vals, ad2 = oracle.assoc_adaptive(p, src, tgt, T, ell, K, precision=oracle.MIRROR)
    x, y = src.xyz.astype(np.float64), moved(T, tgt.xyz)
    n = 0
    capped = 0
    for i in range(src.n):
        a, d2, ok, margin = brute_row(p, x[i], src.feat[i], src.sem[i], y, tgt.feat, tgt.sem, ell, True)
        got = list(cols[i, :counts[i]])
        exp = list(np.nonzero(ok)[0][:K])
        if got != exp:
            diff = set(got) ^ set(exp)
            assert all(margin[j] < 1e-4 for j in diff), (i, diff)
            continue
        for q, j in enumerate(got):
            assert abs(vals[i, q] - a[j]) <= 1e-4 * a[j]
        s = float((a[got] * d2[got]).sum())
        assert abs(ad2[i] - s) <= 2e-4 * max(s, 1e-12)
        n += len(got)
        capped += counts[i] == K
    assert n > 200
    if K < 512:
        assert capped > 30


def fixed_pair_ip(p, rows, ell0, ell):
    """sum of a over a fixed pair set with the geometric factor at ell instead
    of ell0 (the colour and class factors do not depend on ell), FP64."""
    tot = 0.0
    for row in rows:
        for _, a, d2 in row:
            tot += a * math.exp(-d2 / (2 * ell * ell) + d2 / (2 * ell0 * ell0))
    return tot


@pytest.mark.parametrize("ell_max,dl_step", [(1.2, 40.0), (0.5005, -40.0)])
def test_adaptive_lengthscale_step(oracle, ell_max, dl_step):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.is_using_semantics = 0
    p.ell_max = ell_max
    p.dl_step = dl_step
    p.MAX_ITER = 2
    src, tgt, Tgt = synth.planes_spheres_pair(3, 500, 500)
    T0 = np.eye(4, dtype=np.float32)
    To, it, ret, tr = oracle.align_adaptive(p, src, tgt, T0, precision=oracle.POLICY, trace_cap=2)
    assert len(tr) == 2
    ell0 = float(np.float32(p.ell_init))
    K = min(512, p.nearest_neighbors_max)
    x = src.xyz.astype(np.float64)
    y = moved(T0, tgt.xyz)
    A = brute_pairs(p, x, src.feat, None, y, tgt.feat, None, ell0, K, False)
    Axx = brute_pairs(p, x, src.feat, None, x, src.feat, None, ell0, K, False)
    Ayy = brute_pairs(p, y, tgt.feat, None, y, tgt.feat, None, ell0, K, False)
    nA, nxx, nyy = (sum(len(r) for r in R) for R in (A, Axx, Ayy))
    assert tr[0].nnz == nA
    # the closed-form derivative against a central difference of the objective
    terms = ((A, -2.0), (Axx, 1.0), (Ayy, 1.0))
    num = sum(w * a * d2 for R, w in terms for r in R for _, a, d2 in r) / ell0 ** 3
    h = 1e-4

    def obj(e):
        return sum(w * fixed_pair_ip(p, R, ell0, e) for R, w in terms)

    fd = (obj(ell0 + h) - obj(ell0 - h)) / (2 * h)
    assert abs(num - fd) <= 1e-5 * abs(fd)
    dl = num / (nxx + nyy - 2 * nA)
    ell1 = ell0 - dl_step * dl
    if ell1 >= ell_max:
        ell1 = 0.7 * float(np.float32(ell_max))
    ell1 = max(ell1, float(np.float32(p.ell_min)))
    assert abs(tr[1].ell - ell1) <= 4e-7 * ell1, (tr[1].ell, ell1, ell0)
    assert abs(tr[1].ell - ell0) > 1e-4  # the step is resolved


def test_adaptive_step_rule(oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    rng = np.random.default_rng(5)
    seen = set()
    for trial in range(400):
        bcde = rng.normal(size=4) * 10.0 ** rng.uniform(-3, 2, size=4)
        cf = np.array([4 * np.float32(bcde[3]), 3 * np.float32(bcde[2]), 2 * np.float32(bcde[1]),
                       np.float32(bcde[0])], np.float64)
        r = np.roots(cf)
        real = [z.real for z in r if abs(z.imag) <= 1e-9 * max(1.0, abs(z.real)) and z.real > 0]
        want = min(real) if real else float(np.float32(p.min_step))
        want = min(want, 0.8)
        got = oracle.select_step_adaptive(p, bcde)
        assert abs(got - want) <= 1e-6 * max(want, 1e-6), (bcde, got, want, r)
        seen.add("none" if not real else ("cap" if want == 0.8 else "root"))
    assert seen == {"none", "cap", "root"}


@pytest.mark.parametrize("precision", ["MIRROR", "POLICY"])
def test_adaptive_registers(oracle, precision):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.MAX_ITER = 300
    src, tgt, Tgt = synth.planes_spheres_pair(0, 1500, 1500)
    To, it, ret, tr = oracle.align_adaptive(p, src, tgt, np.eye(4, dtype=np.float32),
                                            precision=getattr(oracle, precision), trace_cap=300)
    Ti = np.linalg.inv(Tgt)
    assert rot_dist_rad(To, Ti) < 0.01 * rot_dist_rad(np.eye(4), Ti)
    assert trans_err(To, Ti) < 0.05 * trans_err(np.eye(4), Ti)
    assert it < p.MAX_ITER and ret == 0
    assert len({round(r.ell, 6) for r in tr}) > 5  # the lengthscale moves


def test_adaptive_inner_product_matches_brute_force(oracle, far):
    """Every passing pair at ell_init, no row cap, the target moved by T."""
    p, src, tgt, T = far
    got = oracle.inner_product_adaptive(p, src, tgt, T)
    R, t = T[:3, :3].astype(np.float64), T[:3, 3].astype(np.float64)
    y = tgt.xyz.astype(np.float64) @ R.T + t
    ell = float(np.float32(p.ell_init))
    tot = 0.0
    for i in range(src.n):
        a, _, ok, _ = brute_row(p, src.xyz[i].astype(np.float64), src.feat[i], src.sem[i], y, tgt.feat, tgt.sem,
                                ell, True)
        tot += a[ok].sum()
    want = tot / src.n * 1e6 / tgt.n
    assert want > 0 and abs(got - want) <= 2e-4 * want, (got, want)

