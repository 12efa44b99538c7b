"""Pins the CPU oracle against golden vectors written by library routines
(numpy/scipy, tests/golden/make_golden.py) and against closed forms, brute
force and scripted sequences.  CPU only."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, params_from_yaml
from rkhs_ba_amd import synth
from rkhs_ba_amd.synth import Cloud


def _params(**kw):
    from rkhs_ba_amd import default_params
    p = default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_cubic_roots_vs_numpy(oracle):
    for case in json.load(open(os.path.join(GOLDEN, "cubic_roots.json"))):
        got = oracle.cubic_roots(case["coef"])
        want = np.array(case["re"]) + 1j * np.array(case["im"])
        scale = max(1.0, np.max(np.abs(want)))
        for w in want:  # every numpy root matched by one oracle root
            assert np.min(np.abs(got - w)) <= 1e-6 * scale + 1e-9, (case["coef"], got, want)


def test_step_selection_rule(oracle):
    """Smallest positive real root with |imag| < 1e-5, clamped; no root -> max_step
    (the clamp runs on DBL_MAX, CvoGPU.cu:529-536)."""
    p = _params(min_step=1e-4, max_step=0.5)
    # 4E s^3 + 3D s^2 + 2C s + B with roots 0.2, 0.3, -1 => pick 0.2
    c = np.poly([0.2, 0.3, -1.0])
    bcde = [c[3], c[2] / 2, c[1] / 3, c[0] / 4]
    assert oracle.select_step(p, bcde) == pytest.approx(0.2, rel=1e-9)
    # all roots negative -> no usable root -> max_step
    c = np.poly([-0.2, -0.3, -1.0])
    assert oracle.select_step(p, [c[3], c[2] / 2, c[1] / 3, c[0] / 4]) == pytest.approx(0.5)
    # root above max_step -> max_step; tiny root -> min_step
    c = np.poly([0.9, -0.3, -1.0])
    assert oracle.select_step(p, [c[3], c[2] / 2, c[1] / 3, c[0] / 4]) == pytest.approx(0.5)
    c = np.poly([1e-6, -0.3, -1.0])
    assert oracle.select_step(p, [c[3], c[2] / 2, c[1] / 3, c[0] / 4]) == pytest.approx(1e-4)
    # complex pair with large imaginary part and one negative real root -> max_step
    c = np.poly([0.2 + 0.5j, 0.2 - 0.5j, -1.0]).real
    assert oracle.select_step(p, [c[3], c[2] / 2, c[1] / 3, c[0] / 4]) == pytest.approx(0.5)
    # zero leading coefficient: the companion matrix is undefined -> no root
    assert oracle.select_step(p, [1.0, 1.0, 1.0, 0.0]) == pytest.approx(0.5)


def test_se3_exp_log_vs_scipy(oracle):
    for case in json.load(open(os.path.join(GOLDEN, "se3_exp_log.json"))):
        T = np.array(case["T"]).reshape(3, 4)
        got = oracle.exp_se3_f64(case["xi"], case["dt"])
        assert np.allclose(got, T, atol=1e-12), case
        assert oracle.log_se3_norm(T) == pytest.approx(case["log_norm"], rel=1e-7, abs=1e-12)
        assert np.allclose(oracle.log_se3(T), case["log"], atol=1e-9)
        # float form: agrees to float rounding except the small-angle branch
        f = oracle.exp_sek3_f32(np.float32(case["xi"]), np.float32(case["dt"]))
        th = np.linalg.norm(case["xi"][:3])
        if th >= 1e-6:
            assert np.allclose(f, T, atol=2e-5)


def test_exp_sek3_small_angle_quirk(oracle):
    """LieGroup.cpp:252-254: for |w| < 1e-6 the float form uses J = I, so a pure
    translation twist moves by v, not dt * v."""
    xi = np.array([0, 0, 0, 0.6, 0, 0.8], np.float32)
    f = oracle.exp_sek3_f32(xi, 0.01)
    assert np.allclose(f[:, 3], [0.6, 0, 0.8])
    d = oracle.exp_se3_f64(xi.astype(np.float64), 0.01)
    assert np.allclose(d[:, 3], [0.006, 0, 0.008])


def test_indicator_scripted(oracle):
    """A_sparsity_indicator_ell_update (CvoGPU.cu:547-665), decisions derived by
    hand: the first full start window's last value is also pushed into the end
    window; a stable ratio fires and clears both windows."""
    assert list(oracle.indicator_sequence(2, 0.1, [1, 1, 1, 1, 1, 1])) == [0, 0, 1, 0, 0, 1]
    # growing indicator: ratio 2 at k=2 (no fire), windows slide
    seq = [1, 1, 2, 2, 2, 2]
    # k=2: qs=[1,1] qe=[1,2] ratio 3/2 -> slide: qs=[1,1] qe=[2,2]
    # k=3: ratio 4/2 -> slide qs=[1,2] qe=[2,2]; k=4: 4/3 -> slide qs=[2,2] qe=[2,2]
    # k=5: ratio 1 -> fire
    assert list(oracle.indicator_sequence(2, 0.1, seq)) == [0, 0, 0, 0, 0, 1]
    # W=1: each value fills both windows at once, so the stable ratio fires every call
    assert list(oracle.indicator_sequence(1, 0.5, [3, 3, 3])) == [1, 1, 1]


def test_flow_and_step_coeffs_vs_numpy(oracle):
    for case in json.load(open(os.path.join(GOLDEN, "flow_step.json"))):
        p = _params(c=case["c"], d=case["d"], is_using_range_ell=0)
        x = np.array(case["x"], np.float32)
        z = np.array(case["z"], np.float32)
        counts = np.array(case["counts"], np.int32)
        cols = np.array(case["cols"], np.int32)
        vals = np.array(case["vals"], np.float32)
        f64, u64 = oracle.flow(p, x, z, counts, cols, vals, oracle.FP64)
        assert np.allclose(f64, case["flow"], rtol=1e-12, atol=1e-15)
        f32, u32 = oracle.flow(p, x, z, counts, cols, vals, oracle.MIRROR)
        assert np.allclose(f32, case["flow"], rtol=2e-5, atol=1e-9)
        unit = np.array(case["flow"]) / np.linalg.norm(case["flow"])
        bcde = oracle.step_coeffs(p, case["ell"], x, z, counts, cols, vals, unit, oracle.FP64)
        assert np.allclose(bcde, case["bcde"], rtol=1e-9, atol=1e-12), (bcde, case["bcde"])
        bm = oracle.step_coeffs(p, case["ell"], x, z, counts, cols, vals, unit, oracle.MIRROR)
        assert np.allclose(bm, case["bcde"], rtol=1e-3, atol=1e-6)


def _tiny_clouds(case):
    src = Cloud(np.array(case["x"], np.float32), np.array(case["fx"], np.float32), np.array(case["sx"], np.float32))
    tgt = Cloud(np.array(case["y"], np.float32), np.array(case["fy"], np.float32), np.array(case["sy"], np.float32))
    return src, tgt


def test_assoc_vs_numpy_definition(oracle):
    """Association pairs of the kernel definition (numpy float64 brute force,
    first K by target index); pairs within 1e-6 of a cutoff are exempt."""
    for case in json.load(open(os.path.join(GOLDEN, "assoc_tiny.json"))):
        P = case["params"]
        p = _params(sigma=P["sigma"], sp_thres=P["sp_thres"], c_ell=P["c_ell"], c_sigma=P["c_sigma"],
                    s_ell=P["s_ell"], s_sigma=P["s_sigma"], is_using_intensity=P["is_using_intensity"],
                    is_using_semantics=P["is_using_semantics"])
        src, tgt = _tiny_clouds(case)
        T = np.array(case["T"]).reshape(4, 4)
        K = case["K"]
        for prec in (oracle.MIRROR, oracle.FP64):
            counts, cols, vals, _ = oracle.assoc(p, src, tgt, T, case["ell"], K, precision=prec, brute=True)
            got = {(i, int(cols[i, s])): float(vals[i, s]) for i in range(src.n) for s in range(counts[i])}
            want = {(i, j): (a, m) for i, j, a, m in case["pairs"]}
            for key in set(got) ^ set(want):
                assert key in want and want[key][1] < 1e-6, key
            for key in set(got) & set(want):
                assert got[key] == pytest.approx(want[key][0], rel=2e-6)


def test_assoc_grid_equals_brute(oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, Tgt = synth.planes_spheres_pair(3, 1500, 1800)
    rng = np.random.default_rng(0)
    for ell, K in ((0.5, 512), (0.2, 8), (0.9, 3)):
        T = synth.rigid(synth.rot_axis_angle(rng.normal(size=3), 0.1), rng.normal(size=3) * 0.2)
        a = oracle.assoc(p, src, tgt, T, ell, K, brute=False)
        b = oracle.assoc(p, src, tgt, T, ell, K, brute=True)
        assert (a[0] == b[0]).all() and (a[1] == b[1]).all() and (a[2] == b[2]).all()
        assert a[0].sum() > 0


def test_assoc_cap_keeps_first_k_by_index(oracle):
    """fill_in_A_mat_gpu stops after K stored pairs in ascending target index."""
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, _ = synth.planes_spheres_pair(5, 800, 900)
    full = oracle.assoc(p, src, tgt, np.eye(4), 0.9, 512, brute=True)
    capped = oracle.assoc(p, src, tgt, np.eye(4), 0.9, 4, brute=True)
    assert full[0].max() > 4
    for i in range(src.n):
        k = min(4, full[0][i])
        assert capped[0][i] == k
        assert (capped[1][i, :k] == full[1][i, :k]).all()
        assert (np.diff(full[1][i, :full[0][i]]) > 0).all()


def test_inner_product_invariants(oracle):
    """ip(X, Y, T) with target mapped by T^-1 equals ip(X, T^-1 Y, I) up to float
    rounding of the transform; cos(X, X, I) approx form <= 1 and the exact form is 1."""
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, _ = synth.planes_spheres_pair(1, 2000, 2000)
    T = synth.rigid(synth.rot_axis_angle(np.array([0.3, 1, 0.2]), 0.2), np.array([0.1, -0.2, 0.3]))
    ip1 = oracle.inner_product(p, src, tgt, T, 0.4, oracle.FP64)
    moved = tgt.transformed(np.linalg.inv(T))
    ip2 = oracle.inner_product(p, src, moved, np.eye(4), 0.4, oracle.FP64)
    assert ip1 == pytest.approx(ip2, rel=1e-3)
    ipxx = oracle.inner_product(p, src, src, np.eye(4), 0.4, oracle.FP64)
    assert ipxx > 0


def test_align_converges_to_generating_pose(oracle):
    """Whole loop (mirror precision) recovers the motion of a 2.5k scene to the
    accuracy its sampling allows (independent resamples with 1 cm noise)."""
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.MAX_ITER = 1000
    src, tgt, Tgt = synth.planes_spheres_pair(2, 2500, 2500)
    T, iters, ret, trace = oracle.align(p, src, tgt, np.eye(4), trace_cap=1000)
    from conftest import rot_err_deg, trans_err
    Tinv = np.linalg.inv(Tgt)
    assert ret == 0
    assert rot_err_deg(T, Tinv) < 0.5 and trans_err(T, Tinv) < 0.03
    ells = [r.ell for r in trace]
    assert all(b <= a for a, b in zip(ells, ells[1:]))  # ell never increases


def test_align_disjoint_clouds_returns_minus_one(oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, _ = synth.planes_spheres_pair(2, 300, 300)
    far = tgt.transformed(synth.rigid(np.eye(3), np.array([500.0, 0, 0])))
    T, iters, ret, _ = oracle.align(p, src, far, np.eye(4))
    assert ret == -1 and iters == 0


def test_icosahedral_group_matches_reference_table():
    """The 60 rotations generated by closure equal the reference's table
    (CvoGPU.cpp:281-340) as a set, and the table is a rotation group."""
    path = os.path.join(GOLDEN, "icosahedral_rotations.json")
    table = np.array(json.load(open(path))).reshape(60, 3, 3)
    for R in table:
        assert np.allclose(R @ R.T, np.eye(3), atol=2e-6) and np.linalg.det(R) == pytest.approx(1, abs=1e-5)
    gen = synth.icosahedral_rotations()
    for R in table:
        assert min(np.abs(gen - R).max(axis=(1, 2))) < 2e-6
