"""CPU checks of the oracle's IRLS bundle-adjustment pieces against golden
vectors made by finite differences, scipy.linalg.expm and dense numpy assembly
(tests/golden/ba_pieces.json), plus whole-window behaviour of the oracle's IRLS
loop (src/cvo/IRLS.cpp:148-367)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, params_from_yaml

BA = json.load(open(os.path.join(GOLDEN, "ba_pieces.json")))


def test_se3_plus_matches_expm(oracle):
    # T <- T Exp([rho; phi]), translation first (local_parameterization_se3.hpp:35)
    for case in BA["plus"]:
        T = np.array(case["T"]).reshape(3, 4)
        out = oracle.se3_plus(T, np.array(case["d"]))
        np.testing.assert_allclose(out, np.array(case["out"]).reshape(3, 4), rtol=0, atol=1e-12)


def test_se3_plus_jacobian_matches_finite_differences(oracle):
    # ComputeJacobian at delta = 0 (local_parameterization_se3.hpp:49-83)
    for case in BA["plus_jac"]:
        T = np.array(case["T"]).reshape(3, 4)
        J = oracle.se3_plus_jacobian(T)
        np.testing.assert_allclose(J, np.array(case["J"]).reshape(12, 6), rtol=0, atol=1e-8)


def test_pair_residual_has_reference_jacobian_without_2a(oracle):
    # PairwiseAnalyticalDiffFunctor::Evaluate (IRLS_Cost_CPU.hpp:305-357):
    # the supplied J is d r / d vec(T) divided by 2a
    for case in BA["pair"]:
        r, J1, J2 = oracle.ba_pair_residual(np.array(case["T1"]).reshape(3, 4), np.array(case["T2"]).reshape(3, 4),
                                            np.array(case["p1"]), np.array(case["p2"]), case["a"])
        assert abs(r - case["r"]) <= 1e-12 * max(1.0, abs(case["r"]))
        np.testing.assert_allclose(J1, case["J1"], rtol=1e-7, atol=1e-7)
        np.testing.assert_allclose(J2, case["J2"], rtol=1e-7, atol=1e-7)


def test_normal_equations_match_dense_assembly(oracle):
    ne = BA["normal"]
    poses = np.tile(np.eye(4), (3, 1, 1))
    poses[:, :3, :4] = np.array(ne["poses"]).reshape(3, 3, 4)
    H, g, F = oracle.ba_normal_equations(poses, np.array(ne["pair_frames"]), np.array(ne["p1"], np.float32),
                                         np.array(ne["p2"], np.float32), np.array(ne["a"], np.float32))
    Hg = np.array(ne["H"]).reshape(18, 18)
    np.testing.assert_allclose(H, Hg, rtol=1e-6, atol=1e-7 * np.abs(Hg).max())
    np.testing.assert_allclose(g, ne["g"], rtol=1e-6, atol=1e-7 * np.abs(ne["g"]).max())
    assert abs(F - ne["F"]) <= 1e-10 * max(1.0, abs(ne["F"]))
    # H is symmetric positive semidefinite
    assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() > -1e-9 * np.abs(H).max()


def _window(seed=0, n_frames=4, n_min=1500, n_max=2500):
    from rkhs_ba_amd import synth
    return synth.window_frames(seed, n_frames=n_frames, n_min=n_min, n_max=n_max)


def _rel_errors(P, gts):
    out = []
    for f in range(1, len(P)):
        Ta = np.linalg.inv(P[0]) @ P[f]
        Tb = np.linalg.inv(gts[0]) @ gts[f]
        E = np.linalg.inv(Tb) @ Ta
        out.append((np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1))), np.linalg.norm(E[:3, 3])))
    return np.array(out)


def test_irls_reduces_window_error(oracle):
    p = params_from_yaml("ba_window.yaml")
    p.multiframe_max_iters = 12
    clouds, gts, inits = _window()
    edges = [(i, j) for i in range(4) for j in range(i + 1, 4)]
    P, it, secs = oracle.align_multiframe(p, clouds, inits, [1, 0, 0, 0], edges, threads=8)
    assert it == 12 and len(secs) == 12
    e0, e1 = _rel_errors(inits, gts), _rel_errors(P, gts)
    # the reference's (r, J) moves slowly but steadily toward the truth
    assert e1[:, 0].max() < e0[:, 0].max() and e1[:, 1].max() < e0[:, 1].max()
    np.testing.assert_array_equal(P[0], inits[0])   # held constant


def test_irls_edge_cases(oracle):
    p = params_from_yaml("ba_window.yaml")
    p.multiframe_max_iters = 3
    clouds, gts, inits = _window(n_frames=3)
    # no edges: the loop stops before its first iteration (IRLS.cpp:249-250)
    P, it, _ = oracle.align_multiframe(p, clouds, inits, [0, 0, 0], np.zeros((0, 2), np.int32))
    assert it == 0
    np.testing.assert_array_equal(P[:, :3, :4], inits[:, :3, :4])
    # every frame constant: poses stay, the lengthscale schedule still runs
    P, it, _ = oracle.align_multiframe(p, clouds, inits, [1, 1, 1], [(0, 1), (1, 2)])
    assert it == 3
    np.testing.assert_array_equal(P[:, :3, :4], inits[:, :3, :4])
    # min_nonzeros above any edge's pair count: no edge qualifies
    p.multiframe_min_nonzeros = 10 ** 8
    P, it, _ = oracle.align_multiframe(p, clouds, inits, [1, 0, 0], [(0, 1), (1, 2)])
    assert it == 0


def test_cvo_frame_owns_its_pose():
    # CvoFrame::pose_vec is the frame's own double[12] (CvoFrame.hpp:12-39)
    from rkhs_ba_amd import CvoFrame
    T = np.eye(4)
    f = CvoFrame(None, T)
    f.pose_vec[3] = 2.0
    assert T[0, 3] == 0.0 and f.pose()[0, 3] == 2.0
