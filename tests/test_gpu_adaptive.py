"""GPU adaptive-lengthscale registration (cvo_align_adaptive, the reference's
AdaptiveCvoGPU::align) against the oracle's policy mode (FP64 pose and sums,
the reference's association and float rounding points).

Bars (written in each test):
  * the first 8 iterations step by step: the same stored pair counts, the
    lengthscale within 1e-6 relative, the step within 1e-3 (the root of a
    cubic whose coefficients sum pair values that differ by up to 4e-6: one
    FP32 exp of the summed exponent against three FP64 exps; the trajectory
    doubles such differences every few iterations, as it does a 1e-7 change
    of the start);
  * the end: the run stops on the eps_2 test mid-descent, where a 1e-7 change
    of the start moves the oracle's own end pose by 2.4e-5 rad / 1.3e-4 m and
    its iteration count by ~10, so the GPU end pose is held to 1e-3 of the
    oracle's and both to the generating motion;
  * the lengthscale bound: past ell_max both shrink to 0.7 ell_max.
"""
import os
import sys

import numpy as np
import pytest

from conftest import CONFIGS, params_from_yaml, rot_dist_rad, trans_err
from knn_cases import knn_pair, knn_params
from rkhs_ba_amd import CvoError, CvoGPU, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def actx():
    c = CvoGPU(os.path.join(CONFIGS, "planes_spheres_5k.yaml"))
    yield c
    c.close()


def run_both(ctx, oracle, p, src, tgt, T0, n_trace):
    ctx.write_params(p)
    ctx.set_trace(n_trace)
    ret, T, secs, iters = ctx.align_adaptive(src, tgt, T0)
    tr = ctx.get_trace()
    ctx.set_trace(0)
    To, it_o, ret_o, tr_o = oracle.align_adaptive(p, src, tgt, T0, precision=oracle.POLICY, trace_cap=n_trace)
    return (ret, T, iters, tr), (ret_o, To, it_o, tr_o)


def same_steps(tr, tr_o, n):
    assert len(tr) >= n and len(tr_o) >= n
    for k in range(n):
        a, b = tr[k], tr_o[k]
        assert a.nnz == b.nnz, (k, a.nnz, b.nnz)
        assert abs(a.ell - b.ell) <= 1e-6 * b.ell, (k, a.ell, b.ell)
        assert abs(a.step - b.step) <= 1e-3 * b.step, (k, a.step, b.step)


def test_adaptive_align_matches_oracle(actx, oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, Tgt = synth.planes_spheres_pair(0, 1500, 1500)
    say("adaptive align planes 1.5k")
    (ret, T, iters, tr), (ret_o, To, it_o, tr_o) = run_both(actx, oracle, p, src, tgt, np.eye(4, dtype=np.float32),
                                                            p.MAX_ITER)
    say(f"  gpu {iters} iterations, oracle {it_o}")
    assert ret == ret_o == 0
    same_steps(tr, tr_o, 8)
    assert abs(iters - it_o) <= 30
    assert rot_dist_rad(T, To) <= 1e-3 and trans_err(T, To) <= 1e-3, (rot_dist_rad(T, To), trans_err(T, To))
    Tinv = np.linalg.inv(Tgt)
    assert rot_dist_rad(T, Tinv) < 0.01 * rot_dist_rad(np.eye(4), Tinv)
    assert len({round(r.ell, 6) for r in tr}) > 5  # the lengthscale moves


def test_adaptive_semantics_far_scene(actx, oracle):
    """Colour always, the class factor without a gate, 60 m from the origin."""
    p = knn_params(params_from_yaml("planes_spheres_5k.yaml"))
    p.MAX_ITER = 8
    src, tgt, T = knn_pair(0)
    say("adaptive align far scene, 8 iterations")
    (ret, Tg, iters, tr), (ret_o, To, it_o, tr_o) = run_both(actx, oracle, p, src, tgt, np.eye(4, dtype=np.float32),
                                                             8)
    assert (ret, iters) == (ret_o, it_o)
    same_steps(tr, tr_o, min(len(tr_o), 8))


def test_adaptive_lengthscale_bound(actx, oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    p.ell_max = 0.5005
    p.dl_step = -40.0
    p.MAX_ITER = 6
    src, tgt, Tgt = synth.planes_spheres_pair(3, 500, 500)
    say("adaptive align, ell past ell_max")
    (ret, T, iters, tr), (ret_o, To, it_o, tr_o) = run_both(actx, oracle, p, src, tgt, np.eye(4, dtype=np.float32), 6)
    same_steps(tr, tr_o, min(len(tr_o), 6))
    assert abs(tr[1].ell - 0.7 * np.float32(0.5005)) <= 1e-6


def test_adaptive_errors(actx):
    p = params_from_yaml("planes_spheres_5k.yaml")
    actx.write_params(p)
    src, tgt, _ = synth.planes_spheres_pair(0, 300, 300)
    bare = synth.Cloud(src.xyz, None)
    with pytest.raises(CvoError):
        actx.align_adaptive(bare, tgt, np.eye(4))
    empty = synth.Cloud(np.zeros((0, 3), np.float32), np.zeros((0, 5), np.float32))
    with pytest.raises(CvoError):
        actx.align_adaptive(src, empty, np.eye(4))


def test_adaptive_inner_product_matches_oracle(actx, oracle):
    p = knn_params(params_from_yaml("planes_spheres_5k.yaml"))
    actx.write_params(p)
    src, tgt, T = knn_pair(0)
    for Tq in (T, np.eye(4, dtype=np.float32)):
        say("adaptive inner product")
        got = actx.inner_product_adaptive(src, tgt, Tq)
        want = oracle.inner_product_adaptive(p, src, tgt, Tq)
        assert want > 0 and abs(got - want) <= 1e-5 * want, (got, want)
    empty = synth.Cloud(np.zeros((0, 3), np.float32), np.zeros((0, 5), np.float32))
    assert actx.inner_product_adaptive(src, empty, T) == 0.0

