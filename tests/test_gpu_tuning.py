"""The association-kernel autotuner: every launch shape gives the same pairs.

The knobs (lanes per row per lengthscale range, XCD-contiguous row blocks,
grid cell scale) change only how candidates are enumerated and in which order
a row's pairs are summed, so the bars are: association CSR (sorted by target
id) bit-identical across all 18 fixed settings; inner products equal to 1e-12
relative (FP64 sum order); align poses, with lane counts that switch during
the annealing, within 1e-5 rad / m of the default setting and of the oracle's
policy mode; BA poses within 1e-9 of the oracle after one outer iteration.
"""
import itertools
import os
import sys

import numpy as np
import pytest

from conftest import CONFIGS, params_from_yaml, rot_dist_rad, trans_err
from rkhs_ba_amd import CvoGPU, CvoError, synth

pytestmark = pytest.mark.gpu

SETTINGS = list(itertools.product((16, 8, 4), (0, 1), (1.0, 0.7, 0.5)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture()
def ctx():
    c = CvoGPU(os.path.join(CONFIGS, "planes_spheres_5k.yaml"))
    yield c
    c.close()


def test_settings_give_identical_associations(ctx):
    p = params_from_yaml("street_60k.yaml")
    ctx.write_params(p)
    src, tgt, Tgt = synth.street_pair(4, n=20000)
    for ell in (0.5, 0.1, 0.05):
        ref = None
        for lanes, xcd, scale in SETTINGS:
            say(f"assoc street20k ell={ell} lanes={lanes} xcd={xcd} scale={scale}")
            ctx.set_tuning(lanes, xcd, scale)
            a = ctx.compute_association_gpu(src, tgt, Tgt, ell)
            ip = ctx.inner_product_gpu(src, tgt, Tgt, ell)
            if ref is None:
                ref = (a, ip)
                assert a.nnz > 1000
                continue
            np.testing.assert_array_equal(a.row_ptr, ref[0].row_ptr)
            np.testing.assert_array_equal(a.col, ref[0].col)
            np.testing.assert_array_equal(a.val, ref[0].val)
            assert abs(ip - ref[1]) <= 1e-12 * abs(ref[1]), (ip, ref[1])


def test_settings_align_to_the_same_pose(ctx, oracle):
    p = params_from_yaml("planes_spheres_5k.yaml")
    ctx.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    T0 = np.eye(4, dtype=np.float32)
    T_pol = oracle.align(p, src, tgt, T0, precision=oracle.POLICY)[0]
    poses = {}
    # (lanes, xcd, scale, lanes below split, split, scale below split): 0.15
    # splits the 0.5 -> 0.05 annealing, so lanes and cell scale change inside
    # the captured graph
    for lanes, xcd, scale, small, split, ssmall in [(16, 0, 1.0, 16, 0.0, 1.0), (8, 1, 0.5, 8, 0.0, 0.5),
                                                    (16, 1, 0.7, 8, 0.15, 1.0), (8, 0, 1.0, 16, 0.15, 0.5),
                                                    (16, 1, 0.7, 4, 0.15, 1.0), (4, 0, 1.0, 4, 0.0, 0.7)]:
        say(f"align 5k lanes={lanes}/{small} scale={scale}/{ssmall} split={split} xcd={xcd}")
        ctx.set_tuning(lanes, xcd, scale, lanes_small_ell=small, split_ell=split, cell_scale_small_ell=ssmall)
        ctx.set_trace(p.MAX_ITER)
        ret, T, secs, iters, _ = ctx.align(src, tgt, T0)
        ells = [r.ell for r in ctx.get_trace()]
        if split > 0:
            assert max(ells) > split > min(ells)
        poses[(lanes, xcd, scale, small, split, ssmall)] = T
        assert rot_dist_rad(T, T_pol) <= 1e-5 and trans_err(T, T_pol) <= 1e-5, (lanes, scale, small, split)
    ctx.set_trace(0)
    base = poses[(16, 0, 1.0, 16, 0.0, 1.0)]
    for k, T in poses.items():
        assert rot_dist_rad(T, base) <= 1e-5 and trans_err(T, base) <= 1e-5, k


def test_autotune_measures_caches_and_applies(ctx, tmp_path):
    p = params_from_yaml("street_60k.yaml")
    ctx.write_params(p)
    src, tgt, Tgt = synth.street_pair(4, n=20000)
    ds, dt = ctx.upload(src), ctx.upload(tgt)
    cache = str(tmp_path / "sub" / "tune.json")
    say("autotune street20k (measure)")
    t = ctx.autotune(ds, dt, np.eye(4), cache=cache)
    # 2 block orders x 9 (lanes, scale) settings x 4 lengthscales
    assert t["source"] == "measured" and t["n_tried"] == 72
    assert t["lanes_per_row"] in (4, 8, 16) and t["lanes_small_ell"] in (4, 8, 16) and t["split_ell"] >= 0
    assert t["xcd_remap"] in (0, 1) and t["cell_scale"] in (1.0, 0.7, 0.5)
    assert t["cell_scale_small_ell"] in (1.0, 0.7, 0.5)
    assert t["score_us"] > 0
    assert ctx.get_tuning()["lanes_per_row"] == t["lanes_per_row"]
    key = ctx.tuning_key(ds, dt)
    assert key.startswith("gfx950") and "F5" in key and "K19" in key and key.endswith("|knn0")
    # the K-nearest pass is another kernel: its own entry
    p.is_using_kdtree = 1
    ctx.write_params(p)
    assert ctx.tuning_key(ds, dt) == key[:-1] + "1"
    p.is_using_kdtree = 0
    ctx.write_params(p)
    from rkhs_ba_amd.runtime import tuning_cache_lookup
    e = tuning_cache_lookup(cache, key)
    assert e is not None and e["cell_scale"] == t["cell_scale"]
    # a fresh context reads the cached setting without measuring
    c2 = CvoGPU(os.path.join(CONFIGS, "street_60k.yaml"))
    try:
        d2s, d2t = c2.upload(src), c2.upload(tgt)
        say("autotune street20k (cached)")
        t2 = c2.autotune(d2s, d2t, np.eye(4), cache=cache)
        assert t2["source"] == "cached" and t2["n_tried"] == 72
        keys = ("lanes_per_row", "lanes_small_ell", "split_ell", "xcd_remap", "cell_scale", "cell_scale_small_ell")
        assert all(t2[k] == t[k] for k in keys), (t, t2)
    finally:
        c2.close()
    # the tuned context registers like the default one
    ret, T, secs, iters, _ = ctx.align(ds, dt, np.eye(4))
    assert iters > 0


def test_set_tuning_rejects_bad_values(ctx):
    for bad in [(2, 0, 1.0), (32, 0, 1.0), (16, 2, 1.0), (16, 0, 0.3), (8, 1, 1.5), (8, 1, float("nan"))]:
        with pytest.raises(CvoError):
            ctx.set_tuning(*bad)
    with pytest.raises(CvoError):
        ctx.set_tuning(16, 1, 1.0, lanes_small_ell=2, split_ell=0.1)
    with pytest.raises(CvoError):
        ctx.set_tuning(16, 1, 1.0, lanes_small_ell=8, split_ell=float("inf"))
    with pytest.raises(CvoError):
        ctx.set_tuning(16, 1, 1.0, lanes_small_ell=8, split_ell=0.1, cell_scale_small_ell=0.2)
    t = ctx.get_tuning()
    assert t["source"] == "default" and (t["lanes_per_row"], t["xcd_remap"], t["cell_scale"]) == (16, 1, 1.0)
    ctx.set_tuning(8, 1, 0.5, lanes_small_ell=16, split_ell=0.2)
    t = ctx.get_tuning()
    assert (t["lanes_per_row"], t["lanes_small_ell"], t["xcd_remap"], t["cell_scale"], t["source"]) == (
        8, 16, 1, 0.5, "set")


def test_ba_with_tuned_association_matches_oracle(oracle):
    from test_gpu_ba import all_edges, window
    p = params_from_yaml("ba_window.yaml")
    p.multiframe_max_iters = 1
    clouds, gts, inits = window()
    c = CvoGPU(os.path.join(CONFIGS, "ba_window.yaml"))
    try:
        c.write_params(p)
        # four lanes below the split: the batched BA launch takes eight
        c.set_tuning(16, 1, 0.5, lanes_small_ell=4, split_ell=0.3, cell_scale_small_ell=1.0)
        say("BA window lanes=16/4 scale=0.5/1.0 split=0.3 xcd=1, one outer iteration")
        g = c.align_multiframe(clouds, inits, [1, 0, 0, 0], all_edges(4))
    finally:
        c.close()
    P, it, _ = oracle.align_multiframe(p, clouds, inits, [1, 0, 0, 0], all_edges(4), threads=8)
    assert g["iters"] == it == 1
    np.testing.assert_allclose(g["poses"][:, :3, :4], P[:, :3, :4], rtol=0, atol=1e-9)


def test_four_lane_tuning_in_batched_launches(ctx):
    """A tuning with four lanes per row: single-start align uses it, the
    batched multi-start launch takes eight lanes instead; the starts agree
    with single-start aligns to FP64 summation order (1e-9)."""
    p = params_from_yaml("multistart_20k.yaml")
    p.MAX_ITER = 150
    ctx.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(11, 3000, 3000)
    rng = np.random.default_rng(2)
    Ts = [np.eye(4)]
    for _ in range(2):
        R = synth.rot_axis_angle(rng.normal(size=3), np.deg2rad(rng.uniform(0, 15)))
        Ts.append(synth.rigid(R, rng.normal(size=3) * 0.05))
    Ts = np.array(Ts, np.float32)
    ctx.set_tuning(4, 1, 1.0, lanes_small_ell=4, split_ell=0.0)
    say("multi-start 3 starts with a four-lane tuning")
    out = ctx.align_multistart(src, tgt, Ts, 0.3, mode=0)
    for s in range(len(Ts)):
        say(f"single start {s}, four lanes")
        ret, T, secs, iters, _ = ctx.align(src, tgt, Ts[s])
        assert iters == out["per_start_iters"][s]
        assert rot_dist_rad(T, out["per_start_T"][s]) <= 1e-9 and trans_err(T, out["per_start_T"][s]) <= 1e-9, s
