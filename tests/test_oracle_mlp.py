"""CPU checks of the proposer's FP64 restatement (oracle/mlp_oracle.py): its
backward against central finite differences of its own loss, BF16 rounding
against hand-computed values, and the training-pair generator's labels."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def random_problem(seed, n=2, npts=64):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, npts, 8))
    b = rng.normal(size=(n, npts, 8))
    a[:, :, 3:] = rng.random((n, npts, 5))
    b[:, :, 3:] = rng.random((n, npts, 5))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    tg = np.concatenate([q, rng.normal(size=(n, 3))], 1)
    w = np.concatenate([np.concatenate([rng.uniform(-1, 1, x * y) * np.sqrt(6 / x) * (0.1 if i == 4 else 1.0),
                                        rng.normal(size=y) * 0.05]) for i, (x, y) in enumerate(M.SIZES)])
    return w, a, b, tg


def test_backward_matches_finite_differences():
    w, a, b, tg = random_problem(0)
    L, g = M.gradient(w, a, b, tg, quantize=False)
    rng = np.random.default_rng(1)
    # every layer's weights and biases, sampled
    offs = np.cumsum([0] + [x * y + y for x, y in M.SIZES])
    idx = np.concatenate([rng.choice(np.arange(offs[i], offs[i + 1]), 12, replace=False) for i in range(5)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (M.loss(wp, a, b, tg) - M.loss(wm, a, b, tg)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding():
    # nearest even at the half-way point; 8 significand bits
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-39], np.float32)
    out = M.bf16(vals)
    np.testing.assert_array_equal(out[:5], [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5])
    assert out[5] > 0


def test_parameter_count_and_layout():
    assert M.n_params() == 189696
    w = np.arange(M.n_params(), dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(w)), w)


def test_training_pair_labels_align_the_clouds():
    # target = source-to-target transform: R a + t lands on b's surfaces
    from rkhs_ba_amd import proposer, synth
    a, b, t = proposer.make_pair(3, npts=512)
    R = synth.quat_to_rot(t[:4])
    moved = a[:, :3] @ R.T + t[4:]
    d = np.linalg.norm(moved[:, None, :] - b[None, :, :3], axis=2).min(axis=1)
    assert np.median(d) < 0.5
    assert abs(np.linalg.norm(t[:4]) - 1) < 1e-6 and t[0] >= 0


def test_principal_frame_is_a_signed_eigenbasis():
    rng = np.random.default_rng(5)
    xyz = rng.normal(size=(500, 3)) * [3.0, 2.0, 1.0] + [1, 2, 3]
    xyz[:50] += [8, 0, 0]                      # skew along the first axis
    c, V = M.principal_frame(xyz)
    np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(V) - 1) < 1e-12
    y = xyz - c
    var = np.var(y @ V, axis=0)
    assert var[0] >= var[1] >= var[2] * 0.0
    assert np.sum((y @ V[:, 0]) ** 3) >= 0 and np.sum((y @ V[:, 1]) ** 3) >= 0
    # equivariance: rotating the cloud rotates the frame
    R = M.quat_to_rot(np.array([0.3, -0.5, 0.2, 0.7]))
    c2, V2 = M.principal_frame(xyz @ R.T)
    np.testing.assert_allclose(V2, R @ V, atol=1e-9)
    np.testing.assert_allclose(c2, R @ c, atol=1e-9)
