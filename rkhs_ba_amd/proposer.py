"""Learned warm-start poses for global registration (no counterpart in the
reference): the library's point-set encoder + proposal head (cvo_model_* /
cvo_propose_poses), seeded training pairs, and global registration that refines
the proposals with batched CVO, judges them by the CVO cos angle and falls back
to the plain multi-start grid when they score below a threshold.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from typing import Optional

import numpy as np

from . import synth
from .runtime import CvoGPU, DeviceCloud, _check, lib

K = 8
NPTS = 2048
# Fallback threshold on the approximate cos angle at ell = 0.3 for the
# planes/spheres family (multistart_20k.yaml): over 768 grid starts on 8
# held-out pairs, 95% of converged starts score >= 0.031 and 95% of the others
# <= 0.014.
DEFAULT_MIN_COS = 0.02
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "proposer_k8.npy")


def _setup():
    L = lib()
    if getattr(L, "_mlp_ready", False):
        return L
    P, i32, f = C.c_void_p, C.c_int32, C.c_float
    L.cvo_model_create.argtypes = [P, i32, C.c_uint64, P]
    L.cvo_model_destroy.argtypes = [P]
    L.cvo_model_destroy.restype = None
    L.cvo_model_num_params.restype = C.c_int64
    L.cvo_model_get_weights.argtypes = [P, P, C.c_int64, P]
    L.cvo_model_set_weights.argtypes = [P, P, C.c_int64]
    L.cvo_model_gradient.argtypes = [P, P, i32, i32, P, P, P, P, P]
    L.cvo_model_train_step.argtypes = [P, P, i32, i32, P, P, P, f, P]
    L.cvo_propose_poses.argtypes = [P, P, i32, i32, P, P, P, P, P]
    L.cvo_propose_poses_d.argtypes = [P, P, i32, i32, P, P, P, P, P]
    L.cvo_gemm_bf16.argtypes = [P, i32, i32, i32, P, P, P]
    L._mlp_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def gemm_bf16(cvo: CvoGPU, A, B) -> np.ndarray:
    """The library's BF16 MFMA GEMM on host arrays (C = A B, FP32 result)."""
    L = _setup()
    A, B = _f32(A), _f32(B)
    M, K_ = A.shape
    N = B.shape[1]
    Cm = np.zeros((M, N), np.float32)
    _check(L.cvo_gemm_bf16(cvo._ctx, M, N, K_, A.ctypes.data, B.ctypes.data, Cm.ctypes.data), cvo._ctx)
    return Cm


class PoseProposer:
    """One proposer model in a library context."""

    def __init__(self, cvo: CvoGPU, seed: int = 0, weights: Optional[np.ndarray] = None):
        self._L = _setup()
        self.cvo = cvo
        self._m = C.c_void_p()
        _check(self._L.cvo_model_create(cvo._ctx, K, C.c_uint64(seed), C.byref(self._m)), cvo._ctx)
        if weights is not None:
            self.set_weights(weights)

    @staticmethod
    def num_params() -> int:
        return int(_setup().cvo_model_num_params())

    def close(self):
        if self._m:
            self._L.cvo_model_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.num_params(), np.float32)
        n = C.c_int64()
        _check(self._L.cvo_model_get_weights(self._m, w.ctypes.data, w.size, C.byref(n)), self.cvo._ctx)
        return w

    def set_weights(self, w) -> None:
        w = _f32(w)
        _check(self._L.cvo_model_set_weights(self._m, w.ctypes.data, w.size), self.cvo._ctx)

    def save(self, path: str = WEIGHTS) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.get_weights())

    @classmethod
    def load(cls, cvo: CvoGPU, path: str = WEIGHTS) -> "PoseProposer":
        return cls(cvo, weights=np.load(path, allow_pickle=False))

    def gradient(self, a, b, targets):
        a, b, t = _f32(a), _f32(b), _f32(targets)
        n, npts = a.shape[0], a.shape[1]
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_float()
        _check(self._L.cvo_model_gradient(self.cvo._ctx, self._m, n, npts, a.ctypes.data, b.ctypes.data,
                                          t.ctypes.data, g.ctypes.data, C.byref(loss)), self.cvo._ctx)
        return loss.value, g

    def train_step(self, a, b, targets, lr: float) -> float:
        a, b, t = _f32(a), _f32(b), _f32(targets)
        n, npts = a.shape[0], a.shape[1]
        loss = C.c_float()
        _check(self._L.cvo_model_train_step(self.cvo._ctx, self._m, n, npts, a.ctypes.data, b.ctypes.data,
                                            t.ctypes.data, C.c_float(lr), C.byref(loss)), self.cvo._ctx)
        return loss.value

    def propose(self, a, b):
        """Returns quat_t (n, 8, 7), conf (n, 8), raw head outputs (n, 64).
        a, b: (n, npts, 8) float32 host arrays, or DeviceArray batches
        already in device memory (cvo_propose_poses_d)."""
        from .runtime import DeviceArray
        on_device = isinstance(a, DeviceArray) and isinstance(b, DeviceArray)
        if on_device:
            if len(a.shape) != 3 or a.shape[2] != 8 or a.shape != b.shape:
                raise ValueError("device batches must be two (n, npts, 8) float32 arrays of one shape")
        else:
            a, b = _f32(a), _f32(b)
        n, npts = a.shape[0], a.shape[1]
        qt = np.zeros((n, K, 7), np.float32)
        conf = np.zeros((n, K), np.float32)
        raw = np.zeros((n, 64), np.float32)
        if on_device:
            _check(self._L.cvo_propose_poses_d(self.cvo._ctx, self._m, n, npts, a.ptr, b.ptr,
                                               qt.ctypes.data, conf.ctypes.data, raw.ctypes.data), self.cvo._ctx)
        else:
            _check(self._L.cvo_propose_poses(self.cvo._ctx, self._m, n, npts, a.ctypes.data, b.ctypes.data,
                                             qt.ctypes.data, conf.ctypes.data, raw.ctypes.data), self.cvo._ctx)
        return qt, conf, raw


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------

def cloud_channels(c: synth.Cloud, npts: int = NPTS, seed: int = 0) -> np.ndarray:
    """npts points (downsampled without replacement) as (npts, 8): xyz + 5 features."""
    d = synth.downsample(c, npts, seed) if c.n > npts else c
    feat = d.feat if d.feat is not None else np.zeros((d.n, 5), np.float32)
    return np.concatenate([d.xyz, feat[:, :5]], axis=1).astype(np.float32)


def rot_to_quat(R) -> np.ndarray:
    q = synth.rot_to_quat(np.asarray(R, np.float64))
    return q if q[0] >= 0 else -q


def pose_from_qt(qt) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = synth.quat_to_rot(np.asarray(qt[:4], np.float64))
    T[:3, 3] = qt[4:7]
    return T


def make_pair(seed: int, npts: int = NPTS, trans: float = 2.0):
    """One training pair of the planes/spheres generator: two independent
    samples of a random scene, the second moved by a uniform random rotation
    and a 2 m translation.  Returns (a (npts,8), b (npts,8), target (7,)) with
    the source-to-target transform as the target (align-init convention)."""
    rng = np.random.default_rng([seed, 31])
    scene = synth.PlanesSpheres.make(rng)
    src = scene.sample(np.random.default_rng([seed, 32]), npts)
    tgt0 = scene.sample(np.random.default_rng([seed, 33]), npts)
    R = synth.uniform_rotation(rng)
    t = trans * synth.unit_vector(rng)
    T = synth.rigid(R, t)
    tgt = tgt0.transformed(T)
    target = np.concatenate([rot_to_quat(R), t]).astype(np.float32)
    return cloud_channels(src, npts), cloud_channels(tgt, npts), target


def make_batch(seeds, npts: int = NPTS):
    out = [make_pair(int(s), npts) for s in seeds]
    return (np.stack([o[0] for o in out]), np.stack([o[1] for o in out]), np.stack([o[2] for o in out]))


# ---------------------------------------------------------------------------
# global registration
# ---------------------------------------------------------------------------

def grid_starts(src: synth.Cloud, tgt: synth.Cloud, n: int = 96) -> np.ndarray:
    """The plain multi-start grid: n super-Fibonacci rotations, translations
    matching the centroids (align-init convention)."""
    cx = src.xyz.mean(0).astype(np.float64)
    cy = tgt.xyz.mean(0).astype(np.float64)
    return np.array([synth.rigid(R, cy - R @ cx) for R in synth.so3_grid(n)], np.float32)


def register_global(cvo: CvoGPU, src, tgt, proposer: Optional[PoseProposer] = None, score_ell: float = 0.3,
                    min_cos: Optional[float] = None, grid: int = 96, npts: int = NPTS):
    """Global registration.  With a proposer: its 8 proposals (from 2048-point
    copies of both clouds) run as one batched multi-start, judged by the cos
    angle at score_ell; if the best is below min_cos the plain grid runs too
    and the better result wins.  Without a proposer: the plain grid.
    Returns a dict with T_best (align-output convention), cos, the path taken
    and the seconds spent."""
    t0 = time.perf_counter()
    ds = src if isinstance(src, DeviceCloud) else cvo.upload(src)
    dt = tgt if isinstance(tgt, DeviceCloud) else cvo.upload(tgt)
    out = {"fallback": False}
    if proposer is not None:
        a = cloud_channels(src, npts)[None]
        b = cloud_channels(tgt, npts)[None]
        qt, conf, _ = proposer.propose(a, b)
        starts = np.array([pose_from_qt(q) for q in qt[0]], np.float32)
        r = cvo.align_multistart(ds, dt, starts, score_ell, mode=0)
        out.update(T_best=r["T_best"], cos=float(r["per_start_cos"].max()), proposals=starts, conf=conf[0],
                   per_start_cos=r["per_start_cos"])
        if min_cos is not None and out["cos"] < min_cos:
            g = cvo.align_multistart(ds, dt, grid_starts(src, tgt, grid), score_ell, mode=0)
            out["fallback"] = True
            if float(g["per_start_cos"].max()) > out["cos"]:
                out.update(T_best=g["T_best"], cos=float(g["per_start_cos"].max()))
    else:
        g = cvo.align_multistart(ds, dt, grid_starts(src, tgt, grid), score_ell, mode=0)
        out.update(T_best=g["T_best"], cos=float(g["per_start_cos"].max()), per_start_cos=g["per_start_cos"])
    out["seconds"] = time.perf_counter() - t0
    if ds is not src:
        ds.free()
    if dt is not tgt:
        dt.free()
    return out
