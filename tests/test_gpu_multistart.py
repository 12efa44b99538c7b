"""GPU batched multi-start registration: every start of the batched launch is
bitwise the single-start align from the same init, scores equal the FP64
oracle's inner product, and the grid-scoring mode keeps the reference's rule
(first maximum of function_angle over the rotation grid)."""
import sys

import numpy as np
import pytest

from conftest import params_from_yaml
from rkhs_ba_amd import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def starts_about(T0, n, spread_deg, seed):
    rng = np.random.default_rng(seed)
    out = [np.asarray(T0, np.float64)]
    for _ in range(n - 1):
        R = synth.rot_axis_angle(rng.normal(size=3), np.deg2rad(rng.uniform(0, spread_deg)))
        out.append(np.asarray(T0) @ synth.rigid(R, rng.normal(size=3) * 0.05))
    return np.array(out, np.float32)


def test_batched_starts_equal_single_starts(cvo_gpu, oracle):
    p = params_from_yaml("multistart_20k.yaml")
    p.MAX_ITER = 250
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(11, 3000, 3000)
    Ts = starts_about(np.eye(4), 5, 20.0, 1)
    say("multistart 5 starts")
    out = cvo_gpu.align_multistart(src, tgt, Ts, 0.3, mode=0)
    for s in range(len(Ts)):
        say(f"single start {s}")
        ret, T, secs, iters, _ = cvo_gpu.align(src, tgt, Ts[s])
        assert np.array_equal(T, out["per_start_T"][s]), s
        assert iters == out["per_start_iters"][s]
        # the score is function_angle (approximate) at the final pose
        Tpose = np.linalg.inv(T.astype(np.float64)).astype(np.float32)
        ref = oracle.inner_product(p, src, tgt, Tpose, 0.3, oracle.FP64) / np.sqrt(src.n * tgt.n)
        assert out["per_start_cos"][s] == pytest.approx(ref, rel=1e-4)
    assert out["best_index"] == int(np.argmax(out["per_start_cos"]))
    assert np.array_equal(out["T_best"], out["per_start_T"][out["best_index"]])


def test_grid_scoring_mode_matches_function_angle(cvo_gpu):
    """mode 1 = get_nearest_init_pose: cos of each initial pose, first max."""
    p = params_from_yaml("multistart_20k.yaml")
    p.MAX_ITER = 100
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(12, 2000, 2000, uniform_rot=True)
    rots = synth.icosahedral_rotations()
    Ts = np.array([synth.rigid(R.T, np.zeros(3)) for R in rots], np.float32)
    say("grid60 scoring")
    out = cvo_gpu.align_multistart(src, tgt, Ts, 0.5, mode=1)
    cos_single = np.array([cvo_gpu.function_angle(src, tgt, T, 0.5, True) for T in Ts[:12]])
    assert np.allclose(out["per_start_cos"][:12], cos_single, rtol=1e-9, atol=1e-12)
    assert out["best_index"] == int(np.argmax(out["per_start_cos"]))
    ret, T, secs, iters, _ = cvo_gpu.align(src, tgt, Ts[out["best_index"]])
    assert np.array_equal(T, out["T_best"])
