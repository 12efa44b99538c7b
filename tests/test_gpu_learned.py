"""What the trained proposer gains: its proposals against an untrained network
of the same architecture, and learned global registration (8 proposals refined
by batched CVO, judged by the cos angle, grid fallback) against the plain
96-start grid on 20k-point pairs never seen in training -- the same success at
a fraction of the time."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def err(T, Tgt):
    E = np.asarray(T, np.float64) @ Tgt
    return float(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1))), float(np.linalg.norm(E[:3, 3]))


def test_trained_proposals_beat_untrained(cvo_gpu):
    from rkhs_ba_amd import proposer
    a, b, t = proposer.make_batch(range(9_000_000, 9_000_128))
    trained = proposer.PoseProposer.load(cvo_gpu)
    fresh = proposer.PoseProposer(cvo_gpu, seed=123)
    say("proposals: trained vs untrained")
    out = {}
    for name, p in (("trained", trained), ("untrained", fresh)):
        qt, conf, _ = p.propose(a, b)
        d = np.abs(np.sum(qt[:, :, :4] * t[:, None, :4], axis=2))
        ang = np.degrees(2 * np.arccos(np.clip(d, -1, 1)))
        out[name] = float(np.mean(ang.min(axis=1) < 30.0))
    trained.close()
    fresh.close()
    say(str(out))
    assert out["trained"] >= 0.9 and out["untrained"] <= 0.3


def test_learned_matches_grid_success_at_5x(cvo_gpu):
    from rkhs_ba_amd import read_params_yaml, synth
    from rkhs_ba_amd.proposer import DEFAULT_MIN_COS, PoseProposer, register_global
    from conftest import CONFIGS
    import os
    p = read_params_yaml(os.path.join(CONFIGS, "multistart_20k.yaml"))
    p.MAX_ITER = 1000
    cvo_gpu.write_params(p)
    prop = PoseProposer.load(cvo_gpu)
    ok_g = ok_l = 0
    t_g = t_l = 0.0
    n = 6
    for s in range(n):
        src, tgt, Tgt = synth.planes_spheres_pair(7000 + s, 20000, 20000, uniform_rot=True, trans=2.0)
        say(f"pair {s}: grid")
        g = register_global(cvo_gpu, src, tgt, None)
        say(f"pair {s}: learned")
        l = register_global(cvo_gpu, src, tgt, prop, min_cos=DEFAULT_MIN_COS)
        eg, el = err(g["T_best"], Tgt), err(l["T_best"], Tgt)
        ok_g += eg[0] <= 0.01 and eg[1] <= 0.01
        ok_l += el[0] <= 0.01 and el[1] <= 0.01
        t_g += g["seconds"]
        t_l += l["seconds"]
    prop.close()
    say(f"grid {ok_g}/{n} in {t_g:.2f} s, learned {ok_l}/{n} in {t_l:.2f} s")
    assert ok_l >= ok_g
    assert t_l * 5 <= t_g
