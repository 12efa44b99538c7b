"""GPU learned proposer: the BF16 MFMA GEMM against a host product, forward
outputs and gradients of the library's kernels against the FP64 restatement
(oracle/mlp_oracle.py, BF16 rounding at the same points), training progress,
and proposal decoding.  Tolerance for forward and gradients: 1e-2 relative
(norm-wise), the BF16 bar."""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def prop(cvo_gpu):
    from rkhs_ba_amd.proposer import PoseProposer
    p = PoseProposer(cvo_gpu, seed=7)
    yield p
    p.close()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def batch(seed, n, npts):
    from rkhs_ba_amd import proposer
    return proposer.make_batch(range(1000 * seed, 1000 * seed + n), npts)


def test_gemm_tile_layout(cvo_gpu):
    from rkhs_ba_amd.proposer import gemm_bf16
    rng = np.random.default_rng(0)
    for (m, n, k) in [(16, 16, 16), (100, 70, 40), (257, 129, 33), (64, 256, 128)]:
        say(f"gemm {m}x{n}x{k}")
        A = rng.normal(size=(m, k)).astype(np.float32)
        B = rng.normal(size=(k, n)).astype(np.float32)
        ref = M.bf16(A) @ M.bf16(B)
        out = gemm_bf16(cvo_gpu, A, B)
        assert rel(out, ref) < 1e-6, (m, n, k, rel(out, ref))


def test_forward_matches_oracle(prop):
    a, b, t = batch(1, 3, 256)
    w = prop.get_weights().astype(np.float64)
    say("proposer forward")
    qt, conf, raw = prop.propose(a, b)
    o, _ = M.forward(w, a.astype(np.float64), b.astype(np.float64), quantize=True)
    assert rel(raw, o) < 1e-2
    qo, co, _ = M.propose(w, a.astype(np.float64), b.astype(np.float64))
    np.testing.assert_allclose(np.linalg.norm(qt[:, :, :4], axis=2), 1.0, atol=1e-5)
    assert (qt[:, :, 0] >= 0).all()
    np.testing.assert_allclose(conf.sum(axis=1), 1.0, atol=1e-5)
    assert rel(qt, qo) < 1e-2 and rel(conf, co) < 1e-2


def test_device_batches_give_the_host_result(prop):
    """cvo_propose_poses_d on batches already in device memory: bitwise the
    host-array call."""
    from rkhs_ba_amd import DeviceArray
    a, b, t = batch(2, 5, 512)
    say("proposer host vs device batches")
    qt, conf, raw = prop.propose(a, b)
    da, db = DeviceArray(a), DeviceArray(b)
    qd, cd, rd = prop.propose(da, db)
    assert np.array_equal(qt, qd) and np.array_equal(conf, cd) and np.array_equal(raw, rd)
    with pytest.raises(ValueError):
        prop.propose(da, DeviceArray(b[:4]))


def test_gradient_matches_oracle(prop):
    a, b, t = batch(2, 4, 128)
    w = prop.get_weights().astype(np.float64)
    say("proposer gradient")
    loss, g = prop.gradient(a, b, t)
    Lo, go = M.gradient(w, a.astype(np.float64), b.astype(np.float64), t.astype(np.float64), quantize=True)
    assert abs(loss - Lo) <= 1e-2 * abs(Lo)
    offs = np.cumsum([0] + [x * y + y for x, y in M.SIZES])
    for i in range(5):
        sl = slice(offs[i], offs[i + 1])
        assert rel(g[sl], go[sl]) < 1e-2, (i, rel(g[sl], go[sl]))


def test_gradient_matches_finite_differences(prop):
    # directional derivative of the FP64 loss (central differences) along the
    # library's gradient equals the gradient's norm; the step stays below the
    # loss's kinks (winner switches, ReLU and max-pool boundaries), and 3%
    # covers the BF16 rounding of the library's gradient
    a, b, t = batch(3, 2, 128)
    w = prop.get_weights().astype(np.float64)
    say("proposer gradient vs finite differences")
    _, g = prop.gradient(a, b, t)
    d = g.astype(np.float64) / np.linalg.norm(g)
    eps = 1e-6
    A, B, T = a.astype(np.float64), b.astype(np.float64), t.astype(np.float64)
    fd = (M.loss(w + eps * d, A, B, T) - M.loss(w - eps * d, A, B, T)) / (2 * eps)
    assert abs(fd - np.linalg.norm(g)) <= 0.03 * np.linalg.norm(g), (fd, np.linalg.norm(g))


def test_training_reduces_loss(cvo_gpu):
    from rkhs_ba_amd.proposer import PoseProposer
    p = PoseProposer(cvo_gpu, seed=11)
    a, b, t = batch(4, 16, 256)
    say("proposer training on a fixed batch")
    first = p.train_step(a, b, t, 1e-3)
    for _ in range(60):
        last = p.train_step(a, b, t, 1e-3)
    p.close()
    assert last < 0.7 * first, (first, last)
