"""GPU parity of the two-view path against the CPU oracle (same seeded inputs).

Bars (written in each test):
  * association: identical (row, target) sets except pairs whose oracle value is
    within 1e-6 relative of sp_thres or whose distance is within 1e-6 relative
    of a gate, and a capped row holds exactly the first K passing targets by
    index (tests/assoc_compare.py); stored values within 4e-6 relative (one
    FP32 exp of the summed exponent vs the reference's three FP64 exps
    rounded to float);
  * inner product: 1e-5 relative of the FP64 oracle;
  * align: final pose against the FP32-mirror oracle and the generating pose.
"""
import os
import sys

import numpy as np
import pytest

from assoc_compare import compare_with_oracle
from conftest import rot_dist_rad, CONFIGS, params_from_yaml, rot_err_deg, trans_err
from rkhs_ba_amd import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare_assoc(oracle, params, src, tgt, T, ell, gpu_assoc, K=None, val_rtol=4e-6):
    """The association rule of tests/assoc_compare.py: exact sets except
    near-cutoff pairs; a capped row keeps exactly the first K passing targets
    by index (checked target by target where the row differs)."""
    out = compare_with_oracle(oracle, params, src, tgt, T, ell, gpu_assoc, K=K, val_rtol=val_rtol)
    return out["common"], out["diff"], out["rel"]


def test_association_planes_spheres_fixed_poses(cvo_gpu, oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    rng = np.random.default_rng(1)
    for ell in (0.5, 0.2, 0.05):
        for trial in range(2):
            T = Tgt @ synth.rigid(synth.rot_axis_angle(rng.normal(size=3), rng.uniform(0, 0.05)),
                                  rng.normal(size=3) * 0.05)
            say(f"assoc 5k ell={ell} trial={trial}")
            a = cvo_gpu.compute_association_gpu(src, tgt, T, ell)
            n_common, n_diff, rel = compare_assoc(oracle, p, src, tgt, T, ell, a)
            assert n_common > 100


def test_association_street_semantics(cvo_gpu, oracle):
    """19-class semantic gate on a 20k street scene (the 60k layout, subsampled
    to keep the CPU oracle quick)."""
    p = params_from_yaml("street_60k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.street_pair(4, n=20000)
    for ell in (0.5, 0.1):
        say(f"assoc street20k ell={ell}")
        a = cvo_gpu.compute_association_gpu(src, tgt, Tgt, ell)
        n_common, n_diff, rel = compare_assoc(oracle, p, src, tgt, Tgt, ell, a)
        assert n_common > (100000 if ell == 0.5 else 10000)


def test_association_row_cap_exact_path(cvo_gpu, oracle):
    """A cap far below the rows' pass counts sends rows through the exact
    index-ordered path (first K targets by index, CvoGPU.cu:994-996)."""
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.nearest_neighbors_max = 6
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(6, 3000, 3000)
    say("assoc cap=6")
    a = cvo_gpu.compute_association_gpu(src, tgt, Tgt, 0.9)
    counts = np.diff(a.row_ptr)
    assert counts.max() == 6 and (counts == 6).sum() > 100
    compare_assoc(oracle, p, src, tgt, Tgt, 0.9, a, K=6)


def test_association_row_cap_large_target_semantics(cvo_gpu, oracle):
    """Capped rows against a 70k-point target with the 19-class gate: target
    indices above 2^16 (the selection of the K smallest passing indices takes
    three 8-bit digits), caps of 1, 7 and 100 with thousands of capped rows
    each; the kept pairs are the first K passing targets by index, as the
    oracle keeps them."""
    p = params_from_yaml("street_60k.yaml")
    src, tgt, Tgt = synth.street_pair(4, n=70000)
    src = src.subset(np.arange(0, src.n, 25))
    for K, ell in ((1, 0.5), (7, 0.5), (100, 1.0)):
        p.nearest_neighbors_max = K
        cvo_gpu.write_params(p)
        say(f"assoc cap={K} ell={ell} street 2.8k x 70k")
        a = cvo_gpu.compute_association_gpu(src, tgt, Tgt, ell)
        counts = np.diff(a.row_ptr)
        assert counts.max() == K and (counts == K).sum() > 1000
        assert a.col.max() >= 65536
        compare_assoc(oracle, p, src, tgt, Tgt, ell, a, K=K)


def test_association_geometric_type_and_no_color(cvo_gpu, oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.is_using_intensity = 0
    p.is_using_geometric_type = 1
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(7, 2000, 2000)
    rng = np.random.default_rng(3)
    src.geo_type = rng.uniform(0.05, 1, size=(src.n, 2)).astype(np.float32)
    tgt.geo_type = rng.uniform(0.05, 1, size=(tgt.n, 2)).astype(np.float32)
    say("assoc geo-type")
    a = cvo_gpu.compute_association_gpu(src, tgt, Tgt, 0.4)
    compare_assoc(oracle, p, src, tgt, Tgt, 0.4, a)


def test_inner_product_and_cos_angle(cvo_gpu, oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    for ell in (0.5, 0.1):
        say(f"inner product ell={ell}")
        ip = cvo_gpu.inner_product_gpu(src, tgt, Tgt, ell)
        ref = oracle.inner_product(p, src, tgt, Tgt, ell, oracle.FP64)
        assert ip == pytest.approx(ref, rel=1e-5)
        c = cvo_gpu.function_angle(src, tgt, Tgt, ell, True)
        assert c == pytest.approx(ref / np.sqrt(src.n * tgt.n), rel=1e-5)
    # exact cos of a cloud with itself at identity is 1
    assert cvo_gpu.function_angle(src, src, np.eye(4), 0.3, False) == pytest.approx(1.0, abs=1e-12)
    # empty input gives 0 like the reference
    empty = synth.Cloud(np.zeros((0, 3), np.float32), np.zeros((0, 5), np.float32))
    assert cvo_gpu.function_angle(src, empty, np.eye(4), 0.3) == 0.0


def close_pose(T, To, tol=1e-4):
    return rot_dist_rad(T, To) <= tol and trans_err(T, To) <= tol


def test_align_planes_spheres_matches_oracle(cvo_gpu, oracle):
    # the 5k/5k planes-spheres workload, 2000 iterations.  Default pose
    # arithmetic (FP64): within 1e-5 of the oracle restating exactly that
    # policy, within 1e-4 of the FP32-mirror (reference) and FP64 oracles
    p = params_from_yaml("planes_spheres_5k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    say("align 5k gpu")
    ret, T, secs, iters, _ = cvo_gpu.align(src, tgt, np.eye(4))
    say(f"  iters={iters} secs={secs:.3f}; oracle policy / mirror / fp64")
    for prec, tol in ((oracle.POLICY, 1e-5), (oracle.MIRROR, 1e-4), (oracle.FP64, 1e-4)):
        To, it_o, ret_o, _ = oracle.align(p, src, tgt, np.eye(4), precision=prec)
        assert (ret, iters) == (ret_o, it_o)
        assert close_pose(T, To, tol), (prec, rot_dist_rad(T, To), trans_err(T, To))
    Tinv = np.linalg.inv(Tgt)
    assert rot_err_deg(T, Tinv) < 0.2 and trans_err(T, Tinv) < 0.01


def test_reference_pose_arithmetic_follows_the_mirror_trajectory(cvo_gpu, oracle):
    # with the reference's float pose arithmetic the GPU follows the FP32
    # mirror step by step: poses after 1, 5 and 20 iterations within 1e-6
    # (later, while ell is large, the trajectory amplifies rounding-level
    # differences: a 1e-7 perturbation of the start moves the mirror's own
    # pose by millimetres around iteration 50)
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, _ = synth.planes_spheres_pair(2, 3000, 3000)
    cvo_gpu.set_pose_arithmetic("reference")
    try:
        for n in (1, 5, 20):
            p.MAX_ITER = n
            cvo_gpu.write_params(p)
            say(f"reference pose arithmetic, {n} iterations")
            ret, T, _, it, _ = cvo_gpu.align(src, tgt, np.eye(4))
            To, it_o, ret_o, _ = oracle.align(p, src, tgt, np.eye(4))
            assert (ret, it) == (ret_o, it_o)
            assert close_pose(T, To, 1e-6), (n, rot_dist_rad(T, To), trans_err(T, To))
    finally:
        cvo_gpu.set_pose_arithmetic("fp64")


def test_capped_rows_inside_align_follow_the_mirror(cvo_gpu, oracle):
    # a small nearest_neighbors_max keeps rows over the cap in every
    # iteration (K = min(8, int(1.2 max_row)), CvoGPU.cu:1431), so the capped
    # path (first K passing targets by index, :994-996) feeds the flow and
    # step terms: per-iteration stored pairs and longest row identical to the
    # FP32 mirror, poses within 1e-6 after 1, 5 and 20 iterations
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.nearest_neighbors_max = 8
    src, tgt, _ = synth.planes_spheres_pair(2, 3000, 3000)
    cvo_gpu.set_pose_arithmetic("reference")
    try:
        for n in (1, 5, 20):
            p.MAX_ITER = n
            cvo_gpu.write_params(p)
            cvo_gpu.set_trace(n + 1)
            say(f"capped rows in align, {n} iterations")
            ret, T, _, it, _ = cvo_gpu.align(src, tgt, np.eye(4))
            tr = cvo_gpu.get_trace()
            To, it_o, ret_o, tro = oracle.align(p, src, tgt, np.eye(4), trace_cap=n + 1)
            assert (ret, it) == (ret_o, it_o)
            for k, (g, o) in enumerate(zip(tr[:n], tro[:n])):
                assert (g.nnz, g.max_row, g.num_neighbors) == (o.nnz, o.max_row, o.num_neighbors), k
                assert g.max_row == g.num_neighbors  # rows at the cap in every iteration
            assert close_pose(T, To, 1e-6), (n, rot_dist_rad(T, To), trans_err(T, To))
    finally:
        cvo_gpu.set_trace(0)
        cvo_gpu.set_pose_arithmetic("fp64")


def test_align_street_semantics_matches_oracle(cvo_gpu, oracle):
    # the street family (5 features, 19 classes), 20k/20k, 2000 iterations.
    # Its facades and ground leave translations along them weakly
    # constrained, so rounding-level differences move the final pose: the
    # bound is twice the disagreement between the FP32-mirror and FP64 oracles
    # themselves (at least 1e-4)
    p = params_from_yaml("street_60k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.street_pair(3, n=20000)
    say("align street 20k gpu; oracle policy / mirror / fp64")
    ret, T, _, iters, _ = cvo_gpu.align(src, tgt, np.eye(4))
    Tp, it_p, ret_p, _ = oracle.align(p, src, tgt, np.eye(4), precision=oracle.POLICY)
    Tm, _, _, _ = oracle.align(p, src, tgt, np.eye(4), precision=oracle.MIRROR)
    Tf, _, _, _ = oracle.align(p, src, tgt, np.eye(4), precision=oracle.FP64)
    env_r = max(1e-4, 2 * rot_dist_rad(Tm, Tf))
    env_t = max(1e-4, 2 * trans_err(Tm, Tf))
    assert (ret, iters) == (ret_p, it_p)
    assert rot_dist_rad(T, Tp) <= env_r and trans_err(T, Tp) <= env_t, (rot_dist_rad(T, Tp), trans_err(T, Tp))
    assert rot_dist_rad(T, Tm) <= env_r and trans_err(T, Tm) <= env_t, (rot_dist_rad(T, Tm), trans_err(T, Tm))
    Tinv = np.linalg.inv(Tgt)
    assert rot_err_deg(T, Tinv) < 0.1 and trans_err(T, Tinv) < 0.02


def test_align_is_deterministic_and_graph_equals_eager(cvo_gpu):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.MAX_ITER = 300
    cvo_gpu.write_params(p)
    src, tgt, _ = synth.planes_spheres_pair(1, 3000, 3000)
    say("align graph x2 + eager")
    r1 = cvo_gpu.align(src, tgt, np.eye(4))
    r2 = cvo_gpu.align(src, tgt, np.eye(4))
    cvo_gpu.set_profiling(True)
    try:
        r3 = cvo_gpu.align(src, tgt, np.eye(4))
    finally:
        cvo_gpu.set_profiling(False)
    os.environ["CVO_NO_GRAPH"] = "1"
    try:
        r4 = cvo_gpu.align(src, tgt, np.eye(4))
    finally:
        del os.environ["CVO_NO_GRAPH"]
    assert np.array_equal(r1[1], r2[1]) and np.array_equal(r1[1], r3[1]) and np.array_equal(r1[1], r4[1])
    assert r1[3] == r2[3] == r3[3] == r4[3]


def test_align_trace_and_annealing(cvo_gpu):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.MAX_ITER = 800
    cvo_gpu.write_params(p)
    cvo_gpu.set_trace(800)
    src, tgt, _ = synth.planes_spheres_pair(0)
    say("align with trace")
    cvo_gpu.align(src, tgt, np.eye(4))
    tr = cvo_gpu.get_trace()
    cvo_gpu.set_trace(0)
    assert len(tr) >= 700
    ells = [r.ell for r in tr]
    assert ells[0] == pytest.approx(0.5) and min(ells) < 0.2
    assert all(b <= a for a, b in zip(ells, ells[1:]))
    ks = [r.num_neighbors for r in tr]
    assert ks[0] == 512 and all(k <= 512 for k in ks)
    # K_{k+1} = min(nn_max, int(1.2 max_row_k))  (CvoGPU.cu:1431)
    for a, b in zip(tr, tr[1:]):
        assert b.num_neighbors == min(512, int(a.max_row * 1.2))
    assert all(0 < r.step <= p.max_step + 1e-7 for r in tr)


def test_align_edge_cases(cvo_gpu):
    p = params_from_yaml("planes_spheres_5k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, _ = synth.planes_spheres_pair(2, 500, 500)
    empty = synth.Cloud(np.zeros((0, 3), np.float32), np.zeros((0, 5), np.float32))
    say("align empty")
    ret, T, secs, iters, _ = cvo_gpu.align(src, empty, np.eye(4))
    assert ret == 0 and iters == 0
    res = cvo_gpu.align_result(src, empty, np.eye(4))
    assert res.return_code == -1
    say("align disjoint")
    far = tgt.transformed(synth.rigid(np.eye(3), np.array([500.0, 0, 0])))
    ret, T, secs, iters, _ = cvo_gpu.align(src, far, np.eye(4))
    assert ret == -1 and iters == 0
    # missing features with is_using_intensity raises
    from rkhs_ba_amd import CvoError
    bare = synth.Cloud(src.xyz)
    with pytest.raises(CvoError):
        cvo_gpu.align(bare, tgt, np.eye(4))
