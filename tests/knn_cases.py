"""Synthetic clouds for the K-nearest association tests.

The scene sits about 60 m from the origin, so the range-scaled lengthscale
of the ball association (1 + |x| / 500) differs from the fixed one by ~13%;
the target holds exact duplicates of some of its points (equal distances, so
ties at the K-th place are decided by index); colours, 4-class semantics and
geometric types vary so that every gate rejects some of the nearest targets.
"""
import numpy as np

from rkhs_ba_amd import synth
from rkhs_ba_amd.pointcloud import CvoPointCloud


def _scene(rng, n):
    """Points on three planes and a sphere around (60, 20, 5), with noise."""
    c = np.array([60.0, 20.0, 5.0])
    parts = []
    k = n // 4
    u = rng.uniform(-5, 5, (k, 2))
    parts.append(np.c_[u[:, 0], u[:, 1], np.zeros(k)])            # floor
    u = rng.uniform(-5, 5, (k, 2))
    parts.append(np.c_[np.full(k, -5.0), u[:, 0], u[:, 1] + 5])   # wall x
    u = rng.uniform(-5, 5, (k, 2))
    parts.append(np.c_[u[:, 0], np.full(k, 5.0), u[:, 1] + 5])    # wall y
    m = n - 3 * k
    d = rng.normal(size=(m, 3))
    parts.append(1.5 * d / np.linalg.norm(d, axis=1, keepdims=True) + np.array([1.0, -1.0, 2.0]))
    p = np.concatenate(parts) + c
    return p + rng.normal(scale=0.01, size=p.shape)


def _attributes(rng, xyz):
    ph = xyz @ np.array([[0.7, -0.3, 0.2, 0.1, 0.5], [0.2, 0.6, -0.4, 0.3, 0.1], [0.1, 0.2, 0.5, -0.6, 0.3]])
    feat = 0.5 + 0.25 * np.sin(ph) + rng.normal(scale=0.05, size=ph.shape)
    logits = np.c_[np.sin(xyz[:, 0]), np.cos(xyz[:, 1]), np.sin(xyz[:, 2]), np.cos(xyz[:, 0] + xyz[:, 1])]
    logits = 3 * logits + rng.normal(scale=0.7, size=logits.shape)
    sem = np.exp(logits - logits.max(1, keepdims=True))
    sem /= sem.sum(1, keepdims=True)
    ang = rng.uniform(0, np.pi, len(xyz))
    geo = np.c_[np.cos(ang), np.sin(ang)]
    return feat.astype(np.float32), sem.astype(np.float32), geo.astype(np.float32)


def knn_pair(seed=0, n_src=500, n_tgt=3000, n_dup=300):
    """(source, target, T): target = ground-truth motion of a resampled scene
    plus n_dup duplicated points; T maps the target frame into the source
    frame (the reference's init_guess_transform)."""
    rng = np.random.default_rng(seed)
    xs = _scene(rng, n_src)
    xt0 = _scene(rng, n_tgt)
    # a small motion: the scene's distance from the origin turns 0.3 deg into ~0.3 m
    R = synth.rot_axis_angle(np.array([0.2, 0.3, 1.0]), np.deg2rad(0.3))
    t = np.array([0.1, -0.05, 0.03])
    T = synth.rigid(R, t)
    # target = T * resampled scene (the reference maps it back by T^-1)
    yt = xt0 @ R.T + t
    dup = rng.choice(n_tgt, n_dup, replace=False)
    yt = np.concatenate([yt, yt[dup]])
    fs, ss, gs = _attributes(rng, xs)
    ft, st, gt = _attributes(rng, xt0)
    ft, st, gt = (np.concatenate([a, a[dup]]) for a in (ft, st, gt))
    src = CvoPointCloud.from_arrays(xs.astype(np.float32), fs, ss, gs)
    tgt = CvoPointCloud.from_arrays(yt.astype(np.float32), ft, st, gt)
    return src, tgt, T.astype(np.float32)


def knn_params(base):
    """The planes/spheres parameters with colour, semantics and geometric type on."""
    base.is_using_intensity = 1
    base.is_using_semantics = 1
    base.is_using_geometric_type = 1
    base.sigma = 0.2
    base.c_sigma = 0.8
    base.s_ell = 0.5
    base.s_sigma = 1.0
    return base
