"""The oracle's restatement of CvoGPU::inner_product_cpu (CvoGPU.cpp:114-231)
pinned by an FP64 dense brute force of its definition: the target moved into
the source frame by T^-1, every pair within the squared radius
-2 ell^2 log(sp / sigma^2) at the fixed lengthscale, a = ck k sk (enabled
terms only, no gates), summed where a > sp_thres.  The cases are chosen so
that the plausible mistakes fail: clouds 60 m from the origin (a range-scaled
lengthscale moves the sum by far more than the tolerance), a row cap of 4 (a
capped sum loses most pairs), sigma = 2 (a colour gate would drop pairs whose
product still passes), semantics on and geometry off."""
import numpy as np
import pytest

from conftest import params_from_yaml
from rkhs_ba_amd import synth


def brute(p, src, tgt, T, ell):
    Ti = np.linalg.inv(np.asarray(T, np.float64))
    y = tgt.xyz.astype(np.float64) @ Ti[:3, :3].T + Ti[:3, 3]
    x = src.xyz.astype(np.float64)
    d2 = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    s2 = float(p.sigma) ** 2
    thr = -2.0 * ell * ell * np.log(p.sp_thres / s2)
    a = np.ones_like(d2)
    if p.is_using_geometry:
        a *= s2 * np.exp(-d2 / (2.0 * ell * ell))
    if p.is_using_intensity:
        dc = ((src.feat.astype(np.float64)[:, None, :] - tgt.feat.astype(np.float64)[None, :, :]) ** 2).sum(-1)
        a *= float(p.c_sigma) ** 2 * np.exp(-dc / (2.0 * float(p.c_ell) ** 2))
    if p.is_using_semantics:
        ds = ((src.sem.astype(np.float64)[:, None, :] - tgt.sem.astype(np.float64)[None, :, :]) ** 2).sum(-1)
        a *= float(p.s_sigma) ** 2 * np.exp(-ds / (2.0 * float(p.s_ell) ** 2))
    keep = (d2 < thr) & (a > p.sp_thres)
    return float(a[keep].sum()), int(keep.sum())


def far_pair(seed, n):
    src, tgt, Tgt = synth.planes_spheres_pair(seed, n, n)
    off = synth.rigid(np.eye(3), np.array([60.0, 0.0, 0.0]))
    return src.transformed(off), tgt.transformed(off), off @ Tgt @ np.linalg.inv(off)


@pytest.mark.parametrize("variant", ["default", "sigma2", "no_geometry", "semantics"])
def test_inner_product_cpu_matches_dense_brute_force(oracle, variant):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.nearest_neighbors_max = 4
    p.sigma = 1.0
    src, tgt, T = far_pair(9, 1500)
    if variant == "sigma2":
        p.sigma = 2.0
    elif variant == "no_geometry":
        p.is_using_geometry = 0
    elif variant == "semantics":
        p = params_from_yaml("street_60k.yaml")
        p.nearest_neighbors_max = 4
        p.sigma = 1.0
        src, tgt, T = synth.street_pair(5, n=2500)
    for ell in (1.0, 0.5):
        want, n_want = brute(p, src, tgt, T, ell)
        got, n_got = oracle.inner_product_cpu(p, src, tgt, T, ell)
        got_b, n_b = oracle.inner_product_cpu(p, src, tgt, T, ell, brute=True)
        assert n_want > 6 * src.n  # rows well above the cap of 4
        assert abs(n_got - n_want) <= max(2, n_want // 100000)
        assert got == pytest.approx(want, rel=1e-5)
        assert (got_b, n_b) == (pytest.approx(got, rel=1e-12), n_got)


def test_inner_product_cpu_differs_from_the_gpu_association(oracle):
    """The capped, range-scaled, gated association of inner_product_gpu gives
    another value on the same input (so the two cannot be confused)."""
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.nearest_neighbors_max = 4
    p.sigma = 2.0
    src, tgt, T = far_pair(9, 1500)
    cpu, _ = oracle.inner_product_cpu(p, src, tgt, T, 1.2)
    gpu = oracle.inner_product(p, src, tgt, T, 1.2)
    assert abs(cpu - gpu) > 0.05 * cpu
