"""Seeded synthetic scenes for the registration workloads.

Scenes are generated with numpy's PCG64 generator from an integer seed, so every
run of a test or benchmark sees the same clouds.  All arrays are float32 and
structure-of-arrays (xyz (n,3), feat (n,F), sem (n,K)).

Workloads (names follow what they are):
  * ``planes_spheres``  -- 6 random planes and 3 spheres in a 20 m box, 5-channel
    sinusoidal color features (the 5k two-view case and the 20k multi-start case).
  * ``street``          -- 60k points on a 40x40 m ground plus 9 facades and 15 %
    clutter, 5 color features and 19-class semantic probabilities (the 60k case).
  * ``window_frames``   -- 12 frames of the planes/spheres scene along a smooth
    trajectory, each an independent resample in its sensor frame (the BA window).

The target cloud of a two-view pair is an independent resample of the same scene
mapped by the ground-truth motion ``T_gt``: ``y = R_gt p + t_gt``.  The reference
``align`` returns the map of target coordinates into the source frame, i.e.
``inv(T_gt)``; its init argument is the inverse of that, i.e. ``T_gt`` itself.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class Cloud:
    xyz: np.ndarray
    feat: Optional[np.ndarray] = None
    sem: Optional[np.ndarray] = None
    geo_type: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def transformed(self, T: np.ndarray) -> "Cloud":
        T = np.asarray(T, dtype=np.float64)
        xyz = (self.xyz.astype(np.float64) @ T[:3, :3].T + T[:3, 3]).astype(np.float32)
        return Cloud(xyz, self.feat, self.sem, self.geo_type)

    def subset(self, idx: np.ndarray) -> "Cloud":
        return Cloud(self.xyz[idx], None if self.feat is None else self.feat[idx],
                     None if self.sem is None else self.sem[idx],
                     None if self.geo_type is None else self.geo_type[idx])


def rot_axis_angle(axis: np.ndarray, angle: float) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def rigid(R: np.ndarray, t: np.ndarray) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def unit_vector(rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def uniform_rotation(rng: np.random.Generator) -> np.ndarray:
    """Shoemake's uniform random unit quaternion -> rotation matrix."""
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    return quat_to_rot(np.array([q[3], q[0], q[1], q[2]]))


def quat_to_rot(q: np.ndarray) -> np.ndarray:
    """(w, x, y, z) unit quaternion -> 3x3 rotation."""
    w, x, y, z = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def rot_to_quat(R: np.ndarray) -> np.ndarray:
    """3x3 rotation -> (w, x, y, z) with w >= 0."""
    R = np.asarray(R, dtype=np.float64)
    tr = np.trace(R)
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        q = np.array([0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s])
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = np.array([(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s])
    elif R[1, 1] > R[2, 2]:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = np.array([(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s])
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = np.array([(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s])
    q = q / np.linalg.norm(q)
    return q if q[0] >= 0 else -q


def so3_grid(n: int) -> np.ndarray:
    """Super-Fibonacci spiral samples of SO(3) (Alexa, CVPR 2022): n rotations
    (n,3,3), near-uniformly covering the rotation group."""
    phi = np.sqrt(2.0)
    psi = 1.533751168755204288118041
    out = np.empty((n, 3, 3))
    for i in range(n):
        s = i + 0.5
        r = np.sqrt(s / n)
        R = np.sqrt(1.0 - s / n)
        alpha = 2 * np.pi * s / phi
        beta = 2 * np.pi * s / psi
        q = np.array([r * np.sin(alpha), r * np.cos(alpha), R * np.sin(beta), R * np.cos(beta)])
        out[i] = quat_to_rot(np.array([q[3], q[0], q[1], q[2]]))
    return out


def icosahedral_rotations() -> np.ndarray:
    """The 60 rotations of the icosahedral group in the orientation used by the
    reference's global-initialisation table (src/cvo/CvoGPU.cpp:279-342): a
    3-fold axis along x and a 2-fold axis along y.  Built from the standard
    icosahedron (vertices (0, +-1, +-phi) and cyclic permutations; 5-fold axis
    (0, 1, phi), 3-fold axis (1, 1, 1)) by closure, then conjugated by the frame
    [a3; a2; -(a3 x a2)] with a3 = (1,1,1)/sqrt3 and the perpendicular 2-fold
    axis a2 = (-1/phi, phi, -1)/2 (the mirror-image conjugation selects the
    chirality of that table).  Ordered by (trace, entries).  Returns (60,3,3)."""
    g = (1 + np.sqrt(5)) / 2
    a5 = np.array([0.0, 1.0, g])
    a3 = np.array([1.0, 1.0, 1.0])
    gens = [rot_axis_angle(a5, 2 * np.pi / 5), rot_axis_angle(a3, 2 * np.pi / 3)]
    group = [np.eye(3)]
    frontier = [np.eye(3)]
    while frontier:
        nxt = []
        for M in frontier:
            for G in gens:
                P = G @ M
                if not any(np.allclose(P, Q, atol=1e-9) for Q in group):
                    group.append(P)
                    nxt.append(P)
        frontier = nxt
    assert len(group) == 60, len(group)
    u3 = a3 / np.linalg.norm(a3)
    u2 = np.array([-1.0 / g, g, -1.0]) / 2.0
    Q = np.stack([u3, u2, -np.cross(u3, u2)])
    group = [Q @ M @ Q.T for M in group]
    group.sort(key=lambda M: (-round(np.trace(M), 9), tuple(np.round(M.ravel(), 9))))
    return np.array(group)


# --------------------------------------------------------------------------
# planes + spheres scene
# --------------------------------------------------------------------------

@dataclass
class PlanesSpheres:
    planes: List[Tuple[np.ndarray, np.ndarray, np.ndarray, float, float]] = field(default_factory=list)
    spheres: List[Tuple[np.ndarray, float]] = field(default_factory=list)
    omegas: np.ndarray = None
    phases: np.ndarray = None

    @staticmethod
    def make(rng: np.random.Generator, box: float = 10.0) -> "PlanesSpheres":
        sc = PlanesSpheres()
        for _ in range(6):
            c = rng.uniform(-box, box, size=3) * 0.6
            nrm = unit_vector(rng)
            u = np.cross(nrm, unit_vector(rng))
            u /= np.linalg.norm(u)
            v = np.cross(nrm, u)
            hu, hv = rng.uniform(2.0, 6.0, size=2)
            sc.planes.append((c, u, v, float(hu), float(hv)))
        for _ in range(3):
            c = rng.uniform(-box, box, size=3) * 0.6
            sc.spheres.append((c, float(rng.uniform(0.5, 2.0))))
        dirs = np.stack([unit_vector(rng) for _ in range(5)])
        sc.omegas = dirs * rng.uniform(0.3, 1.2, size=(5, 1))
        sc.phases = rng.uniform(0, 2 * np.pi, size=5)
        return sc

    def sample(self, rng: np.random.Generator, n: int, noise: float = 0.01) -> Cloud:
        areas = [4 * hu * hv for (_, _, _, hu, hv) in self.planes] + [4 * np.pi * r * r for (_, r) in self.spheres]
        areas = np.array(areas)
        counts = rng.multinomial(n, areas / areas.sum())
        pts = []
        for k, cnt in enumerate(counts):
            if cnt == 0:
                continue
            if k < len(self.planes):
                c, u, v, hu, hv = self.planes[k]
                s = rng.uniform(-1, 1, size=(cnt, 2))
                pts.append(c + s[:, :1] * hu * u + s[:, 1:] * hv * v)
            else:
                c, r = self.spheres[k - len(self.planes)]
                d = rng.normal(size=(cnt, 3))
                d /= np.linalg.norm(d, axis=1, keepdims=True)
                pts.append(c + r * d)
        P = np.concatenate(pts, axis=0)
        P = P[rng.permutation(P.shape[0])]
        feat = 0.5 + 0.5 * np.sin(P @ self.omegas.T + self.phases)
        xyz = P + rng.normal(scale=noise, size=P.shape)
        return Cloud(xyz.astype(np.float32), feat.astype(np.float32))


def planes_spheres_pair(seed: int, n_src: int = 5000, n_tgt: int = 5000, angle_deg: float = 10.0,
                        trans: float = 0.3, uniform_rot: bool = False):
    """Returns (source, target, T_gt) with target = T_gt * resample."""
    rng = np.random.default_rng([seed, 1])
    scene = PlanesSpheres.make(rng)
    src = scene.sample(np.random.default_rng([seed, 2]), n_src)
    tgt0 = scene.sample(np.random.default_rng([seed, 3]), n_tgt)
    rr = np.random.default_rng([seed, 4])
    if uniform_rot:
        R = uniform_rotation(rr)
    else:
        R = rot_axis_angle(unit_vector(rr), np.deg2rad(angle_deg))
    t = trans * unit_vector(rr)
    T_gt = rigid(R, t)
    return src, tgt0.transformed(T_gt), T_gt


# --------------------------------------------------------------------------
# street scene: ground + facades + clutter, 19 semantic classes
# --------------------------------------------------------------------------

@dataclass
class Street:
    segments: List[Tuple[np.ndarray, np.ndarray, np.ndarray, float, float, int]]
    omegas: np.ndarray
    phases: np.ndarray
    n_classes: int = 19

    @staticmethod
    def make(rng: np.random.Generator) -> "Street":
        segs = [(np.array([0.0, 0.0, 0.0]), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), 20.0, 20.0, 0)]
        classes = rng.choice(np.arange(1, 19), size=9, replace=False)
        for k in range(9):
            w = rng.uniform(5, 15)
            h = rng.uniform(3, 8)
            yaw = rng.uniform(0, 2 * np.pi)
            u = np.array([np.cos(yaw), np.sin(yaw), 0.0])
            v = np.array([0.0, 0.0, 1.0])
            c = np.array([rng.uniform(-15, 15), rng.uniform(-15, 15), h / 2])
            segs.append((c, u, v, w / 2, h / 2, int(classes[k])))
        dirs = np.stack([unit_vector(rng) for _ in range(5)])
        om = dirs * rng.uniform(0.3, 1.2, size=(5, 1))
        ph = rng.uniform(0, 2 * np.pi, size=5)
        return Street(segs, om, ph)

    def sample(self, rng: np.random.Generator, n: int, noise: float = 0.01) -> Cloud:
        n_seg = int(round(0.85 * n))
        n_clut = n - n_seg
        areas = np.array([4 * hu * hv for (_, _, _, hu, hv, _) in self.segments])
        counts = rng.multinomial(n_seg, areas / areas.sum())
        pts, cls = [], []
        for (c, u, v, hu, hv, k), cnt in zip(self.segments, counts):
            s = rng.uniform(-1, 1, size=(cnt, 2))
            pts.append(c + s[:, :1] * hu * u + s[:, 1:] * hv * v)
            cls.append(np.full(cnt, k))
        clut = np.stack([rng.uniform(-20, 20, n_clut), rng.uniform(-20, 20, n_clut), rng.uniform(0, 8, n_clut)], axis=1)
        pts.append(clut)
        cls.append(np.full(n_clut, -1))
        P = np.concatenate(pts)
        C = np.concatenate(cls)
        perm = rng.permutation(P.shape[0])
        P, C = P[perm], C[perm]
        feat = 0.5 + 0.5 * np.sin(P @ self.omegas.T + self.phases)
        z = rng.normal(size=(P.shape[0], self.n_classes))
        has = C >= 0
        z[np.nonzero(has)[0], C[has]] += 4.0
        z -= z.max(axis=1, keepdims=True)
        e = np.exp(z)
        sem = e / e.sum(axis=1, keepdims=True)
        xyz = P + rng.normal(scale=noise, size=P.shape)
        return Cloud(xyz.astype(np.float32), feat.astype(np.float32), sem.astype(np.float32))


def street_pair(seed: int, n: int = 60000, angle_deg: float = 5.0, trans: float = 1.0):
    rng = np.random.default_rng([seed, 11])
    scene = Street.make(rng)
    src = scene.sample(np.random.default_rng([seed, 12]), n)
    tgt0 = scene.sample(np.random.default_rng([seed, 13]), n)
    rr = np.random.default_rng([seed, 14])
    R = rot_axis_angle(unit_vector(rr), np.deg2rad(angle_deg))
    t = trans * unit_vector(rr)
    T_gt = rigid(R, t)
    return src, tgt0.transformed(T_gt), T_gt


# --------------------------------------------------------------------------
# BA window
# --------------------------------------------------------------------------

def _catmull_rom(P: np.ndarray, s: float) -> Tuple[np.ndarray, np.ndarray]:
    """Position and tangent on a Catmull-Rom spline through control points P."""
    n = P.shape[0] - 1
    u = np.clip(s, 0, 1) * n
    i = min(int(u), n - 1)
    t = u - i
    p0 = P[max(i - 1, 0)]
    p1, p2 = P[i], P[i + 1]
    p3 = P[min(i + 2, n)]
    pos = 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t + (-p0 + 3 * p1 - 3 * p2 + p3) * t ** 3)
    tan = 0.5 * ((-p0 + p2) + 2 * (2 * p0 - 5 * p1 + 4 * p2 - p3) * t + 3 * (-p0 + 3 * p1 - 3 * p2 + p3) * t * t)
    return pos, tan


def window_frames(seed: int, n_frames: int = 12, n_min: int = 5000, n_max: int = 40000,
                  rot_noise_deg: float = 3.0, trans_noise: float = 0.2):
    """12 frames of the planes/spheres scene.  Returns (clouds_in_sensor_frame,
    gt_poses (n,4,4) world<-sensor, init_poses (n,4,4)); frame 0 init = gt."""
    rng = np.random.default_rng([seed, 21])
    scene = PlanesSpheres.make(rng)
    ctrl = rng.uniform(-4, 4, size=(4, 3))
    clouds, gts, inits = [], [], []
    for f in range(n_frames):
        s = f / max(1, n_frames - 1)
        pos, tan = _catmull_rom(ctrl, s)
        yaw = np.arctan2(tan[1], tan[0])
        R = rot_axis_angle(np.array([0, 0, 1.0]), yaw)
        Tw = rigid(R, pos)
        nf = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        world = scene.sample(np.random.default_rng([seed, 100 + f]), nf)
        clouds.append(world.transformed(np.linalg.inv(Tw)))
        gts.append(Tw)
        if f == 0:
            inits.append(Tw.copy())
        else:
            dR = rot_axis_angle(unit_vector(rng), np.deg2rad(rot_noise_deg))
            dt = trans_noise * unit_vector(rng)
            inits.append(Tw @ rigid(dR, dt))
    return clouds, np.array(gts), np.array(inits)


def downsample(c: Cloud, n: int, seed: int = 0) -> Cloud:
    rng = np.random.default_rng(seed)
    idx = rng.choice(c.n, size=min(n, c.n), replace=False)
    idx.sort()
    return c.subset(idx)
