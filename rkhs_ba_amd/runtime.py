"""Python host mirror of the reference ``cvo::CvoGPU`` interface over the C ABI.

The methods keep the reference's names, argument order and meaning
(include/UnifiedCvo/cvo/CvoGPU.hpp:43-274).  All compute runs in
``librkhs_ba_amd.so`` (HIP, gfx950); there is no CPU fallback: when the library
or a GPU is missing the calls raise ``CvoError``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from ._abi import (Assoc, CvoParams, HeldView, IterRecord, KernelStats, KSTAT_NAMES, Tuning, mat16)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librkhs_ba_amd.so")
_lib = None

STATUS = {0: "ok", 1: "invalid argument", 2: "empty point cloud", 3: "file could not be opened",
          4: "YAML value could not be parsed", 5: "HIP runtime error", 6: "no HIP device",
          7: "out of device memory", 8: "unsupported option", 9: "size beyond a documented limit"}


class CvoError(RuntimeError):
    def __init__(self, status: int, msg: str = ""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}{': ' + msg if msg else ''}")


def lib() -> C.CDLL:
    """Loads librkhs_ba_amd.so (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CvoError(5, f"{LIB_PATH} is missing; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P, i32, f, d = C.c_void_p, C.c_int32, C.c_float, C.c_double
        L.cvo_status_string.restype = C.c_char_p
        L.cvo_last_error.restype = C.c_char_p
        L.cvo_last_error.argtypes = [P]
        L.cvo_version.restype = C.c_char_p
        L.cvo_cloud_size.restype = i32
        L.cvo_cloud_size.argtypes = [P]
        L.cvo_destroy.argtypes = [P]
        L.cvo_destroy.restype = None
        L.cvo_cloud_free.argtypes = [P, P]
        L.cvo_cloud_free.restype = None
        L.cvo_assoc_free.argtypes = [P]
        L.cvo_assoc_free.restype = None
        L.cvo_inner_product.argtypes = [P, P, P, P, f, P]
        L.cvo_inner_product_d.argtypes = [P, P, P, P, f, P]
        L.cvo_function_angle.argtypes = [P, P, P, P, f, i32, P]
        L.cvo_compute_association.argtypes = [P, P, P, P, f, P]
        L.cvo_compute_association_nonisotropic.argtypes = [P, P, P, P, P, P]
        L.cvo_compute_association_knn.argtypes = [P, P, P, P, f, i32, P]
        L.cvo_align_adaptive.argtypes = [P, P, P, P, P, P, P, P]
        L.cvo_align_adaptive_d.argtypes = [P, P, P, P, P, P, P, P]
        L.cvo_inner_product_adaptive.argtypes = [P, P, P, P, P]
        L.cvo_align.argtypes = [P, P, P, P, P, P, P, P, P]
        L.cvo_align_d.argtypes = [P, P, P, P, P, P, P, P, P]
        L.cvo_align_multistart.argtypes = [P, P, P, i32, P, f, i32, P, P, P, P, P, P]
        L.cvo_align_multiframe.argtypes = [P, i32, P, P, P, i32, P, P, P, P, P, i32]
        L.cvo_align_s2t.argtypes = [P, P, P, P, P, P, P, P, P]
        L.cvo_align_s2t_d.argtypes = [P, P, P, P, P, P, P, P, P]
        L.cvo_get_nearest_init_pose.argtypes = [P, P, P, f, i32, P, P]
        L.cvo_get_nearest_init_pose_d.argtypes = [P, P, P, f, i32, P, P]
        L.cvo_icosahedral_rotations.argtypes = [P]
        L.cvo_inner_product_cpu.argtypes = [P, P, P, P, f, P]
        L.cvo_function_angle_ex.argtypes = [P, P, P, P, f, i32, i32, P]
        L.cvo_inner_product_d_noalloc.argtypes = [P, P, P, P, f, P]
        L.cvo_set_pose_arithmetic.argtypes = [P, i32]
        L.cvo_workspace_bytes.argtypes = [P, P]
        L.cvo_set_workspace_limit.argtypes = [P, C.c_uint64]
        L.cvo_release_workspace.argtypes = [P]
        L.cvo_get_tuning.argtypes = [P, P]
        L.cvo_set_tuning.argtypes = [P, P]
        L.cvo_autotune.argtypes = [P, P, P, P, C.c_char_p, P]
        L.cvo_autotune_d.argtypes = [P, P, P, P, C.c_char_p, P]
        L.cvo_tuning_key.argtypes = [P, P, P, C.c_char_p, i32]
        L.cvo_tuning_cache_lookup.argtypes = [C.c_char_p, C.c_char_p, P]
        L.cvo_tuning_cache_store.argtypes = [C.c_char_p, C.c_char_p, P]
        _lib = L
    return _lib


def _check(st: int, ctx=None):
    if st != 0:
        msg = ""
        if ctx:
            msg = lib().cvo_last_error(ctx).decode(errors="replace")
        raise CvoError(st, msg)


def default_tuning_cache() -> str:
    """Tuning cache file: $RKHS_BA_AMD_TUNE_CACHE, else
    ~/.cache/rkhs_ba_amd/autotune.json."""
    env = os.environ.get("RKHS_BA_AMD_TUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "rkhs_ba_amd", "autotune.json")


def tuning_cache_lookup(path: str, key: str) -> Optional[dict]:
    """Entry of a tuning cache file (no device needed); None when absent."""
    t = Tuning()
    st = lib().cvo_tuning_cache_lookup(path.encode(), key.encode(), C.byref(t))
    if st == 3:
        return None
    _check(st)
    return t.as_dict()


def _tuning(lanes_per_row, xcd_remap, cell_scale, lanes_small_ell, split_ell, cell_scale_small_ell, source=1,
            score_us=0.0, n_tried=0) -> Tuning:
    small = lanes_per_row if lanes_small_ell is None else lanes_small_ell
    scale_small = cell_scale if cell_scale_small_ell is None else cell_scale_small_ell
    return Tuning(int(lanes_per_row), int(small), float(split_ell), int(xcd_remap), float(cell_scale),
                  float(scale_small), int(source), float(score_us), int(n_tried))


def tuning_cache_store(path: str, key: str, lanes_per_row: int, xcd_remap: int, cell_scale: float,
                       lanes_small_ell: Optional[int] = None, split_ell: float = 0.0,
                       cell_scale_small_ell: Optional[float] = None, score_us: float = 0.0, n_tried: int = 0) -> None:
    t = _tuning(lanes_per_row, xcd_remap, cell_scale, lanes_small_ell, split_ell, cell_scale_small_ell,
                score_us=score_us, n_tried=n_tried)
    _check(lib().cvo_tuning_cache_store(path.encode(), key.encode(), C.byref(t)))


def default_params() -> CvoParams:
    """CvoParams::CvoParams() defaults (CvoParams.hpp:80-136)."""
    p = CvoParams()
    _check(lib().cvo_params_default(C.byref(p)))
    return p


def read_params_yaml(path: str, base: Optional[CvoParams] = None) -> CvoParams:
    """read_CvoParams_yaml (CvoParams.hpp:203-323) on top of ``base`` (defaults)."""
    p = default_params() if base is None else base.copy()
    _check(lib().cvo_params_read_yaml(os.fsencode(path), C.byref(p)))
    return p


@dataclass
class Association:
    """cvo::Association (Association.hpp:7-14) as CSR: row_ptr (N+1), col, val.
    source_inliers / target_inliers follow gpu_association_to_cpu
    (CvoGPU_impl.cu:370-431) without its duplicate-order nondeterminism."""
    rows: int
    cols: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1]) if self.row_ptr.size else 0

    @property
    def source_inliers(self) -> np.ndarray:
        return np.nonzero(np.diff(self.row_ptr) > 0)[0]

    @property
    def target_inliers(self) -> np.ndarray:
        return self.col.copy()

    def pairs(self):
        """(row, col, val) triplets."""
        rows = np.repeat(np.arange(self.rows), np.diff(self.row_ptr))
        return rows, self.col, self.val


@dataclass
class CvoResultInfo:
    """CvoResultInfo (CvoGPU.hpp:34-40)."""
    T_s2t: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    num_iters: int = 0
    registration_seconds: float = 0.0
    return_code: int = 0
    association: Optional[Association] = None


def _take_assoc(a: Assoc) -> Association:
    n = a.rows
    nnz = int(a.nnz)
    if a.row_ptr:
        row_ptr = np.ctypeslib.as_array(a.row_ptr, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(a.col, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(a.val, shape=(max(nnz, 1),))[:nnz].copy()
    else:
        row_ptr = np.zeros(n + 1, np.int64)
        col = np.zeros(0, np.int32)
        val = np.zeros(0, np.float32)
    lib().cvo_assoc_free(C.byref(a))
    return Association(n, int(a.cols), row_ptr, col, val)


_hip_rt = None


def _hip():
    """The HIP runtime the library is linked against (already loaded with it;
    a second HIP runtime in the process, such as the one a framework wheel
    bundles, cannot share the device with it)."""
    global _hip_rt
    if _hip_rt is None:
        lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip_rt = h
    return _hip_rt


class DeviceArray:
    """A float32 array copied into device memory through the library's own
    HIP runtime: the input of the entry points that take device pointers
    (cvo_propose_poses_d).  .ptr is the device address."""

    def __init__(self, arr):
        a = np.ascontiguousarray(arr, np.float32)
        self.shape, self.dtype, self.nbytes = a.shape, a.dtype, a.nbytes
        h = _hip()
        p = C.c_void_p()
        if h.hipMalloc(C.byref(p), max(1, a.nbytes)) != 0:
            raise CvoError(7, "hipMalloc failed")
        self.ptr = p.value
        if h.hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 1) != 0:  # host to device
            self.free()
            raise CvoError(5, "hipMemcpy failed")

    def free(self):
        if getattr(self, "ptr", None):
            _hip().hipFree(C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCloud:
    """A cloud resident in device memory (cvo_cloud_upload)."""

    def __init__(self, owner: "CvoGPU", cloud):
        self._owner = owner
        self._held = HeldView(cloud)
        self.handle = C.c_void_p()
        _check(lib().cvo_cloud_upload(owner._ctx, self._held.ref, C.byref(self.handle)), owner._ctx)
        self.n = self._held.view.n

    def free(self):
        if self.handle and self._owner._ctx:
            lib().cvo_cloud_free(self._owner._ctx, self.handle)
        self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CvoFrame:
    """cvo::CvoFrame (CvoFrame.hpp:12-39): a sensor-frame cloud and its pose
    (12 doubles, row-major 3x4, world <- sensor) updated in place by
    CvoGPU.align_frames."""

    def __init__(self, points, pose_vec=None):
        self.points = points
        pv = np.array(np.eye(4) if pose_vec is None else pose_vec, np.float64)  # own copy, like double[12]
        if pv.ndim == 2:
            pv = pv[:3, :4]
        self.pose_vec = np.ascontiguousarray(pv.reshape(12)).copy()

    def pose(self) -> np.ndarray:
        T = np.eye(4)
        T[:3, :4] = self.pose_vec.reshape(3, 4)
        return T


class CvoGPU:
    """cvo::CvoGPU (CvoGPU.hpp:43-274) on one MI355X."""

    def __init__(self, param_file: Optional[str] = None, params: Optional[CvoParams] = None):
        self._ctx = C.c_void_p()
        if params is not None:
            _check(lib().cvo_create_from_params(C.byref(params), C.byref(self._ctx)))
        else:
            _check(lib().cvo_create(os.fsencode(param_file) if param_file else None, C.byref(self._ctx)))

    def close(self):
        if self._ctx:
            lib().cvo_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- params -------------------------------------------------------------
    def get_params(self) -> CvoParams:
        p = CvoParams()
        _check(lib().cvo_get_params(self._ctx, C.byref(p)))
        return p

    def write_params(self, p: CvoParams) -> None:
        _check(lib().cvo_set_params(self._ctx, C.byref(p)), self._ctx)

    # -- device clouds ------------------------------------------------------
    def upload(self, cloud) -> DeviceCloud:
        return DeviceCloud(self, cloud)

    # -- two-view -----------------------------------------------------------
    def align(self, source, target, init_guess_transform, association: bool = False):
        """int align(source, target, init_guess_T, transform, association,
        registration_seconds, num_iters) (CvoGPU.cu:1516-1563).  Returns
        (ret, transform 4x4, seconds, iters, association-or-None)."""
        T0 = mat16(init_guess_transform)
        Tout = np.zeros(16, np.float32)
        secs = C.c_double()
        iters = C.c_int32()
        ret = C.c_int32()
        a = Assoc()
        if isinstance(source, DeviceCloud):
            st = lib().cvo_align_d(self._ctx, source.handle, target.handle, T0.ctypes.data, Tout.ctypes.data,
                                   C.byref(a) if association else None, C.byref(secs), C.byref(iters), C.byref(ret))
        else:
            hs, ht = HeldView(source), HeldView(target)
            st = lib().cvo_align(self._ctx, hs.ref, ht.ref, T0.ctypes.data, Tout.ctypes.data,
                                 C.byref(a) if association else None, C.byref(secs), C.byref(iters), C.byref(ret))
        if st == 2:  # empty input: the reference returns 0 and leaves transform
            return 0, np.eye(4, dtype=np.float32), 0.0, 0, None
        _check(st, self._ctx)
        assoc = _take_assoc(a) if (association and a.row_ptr) else None
        return ret.value, Tout.reshape(4, 4), secs.value, iters.value, assoc

    def align_adaptive(self, source, target, init_guess_transform):
        """AdaptiveCvoGPU::align (AdaptiveCvoGPU.cu:1077-1248): the lengthscale
        follows its gradient from the self inner products.  Returns (ret,
        transform 4x4, seconds, iters)."""
        T0 = mat16(init_guess_transform)
        Tout = np.zeros(16, np.float32)
        secs = C.c_double()
        iters = C.c_int32()
        ret = C.c_int32()
        if isinstance(source, DeviceCloud):
            st = lib().cvo_align_adaptive_d(self._ctx, source.handle, target.handle, T0.ctypes.data,
                                            Tout.ctypes.data, C.byref(secs), C.byref(iters), C.byref(ret))
        else:
            hs, ht = HeldView(source), HeldView(target)
            st = lib().cvo_align_adaptive(self._ctx, hs.ref, ht.ref, T0.ctypes.data, Tout.ctypes.data,
                                          C.byref(secs), C.byref(iters), C.byref(ret))
        _check(st, self._ctx)
        return ret.value, Tout.reshape(4, 4), secs.value, iters.value

    def inner_product_adaptive(self, source, target, s2t_frame_transform) -> float:
        """AdaptiveCvoGPU::inner_product (AdaptiveCvoGPU.cu:1338-1371): the
        target moved by the transform, the adaptive association at ell_init
        without a row cap, sum(a) / N * 1e6 / M."""
        T0 = mat16(s2t_frame_transform)
        hs, ht = HeldView(source), HeldView(target)
        out = C.c_float()
        _check(lib().cvo_inner_product_adaptive(self._ctx, hs.ref, ht.ref, T0.ctypes.data, C.byref(out)), self._ctx)
        return out.value

    def align_result(self, source, target, init_guess_T_s2t, is_returning_association: bool = False) -> CvoResultInfo:
        """CvoResultInfo align(source, target, T_init_s2t, is_returning_association)
        (CvoGPU.cu:1566-1611): starts from the inverse of its init, or with
        is_global_angle_registration from get_nearest_init_pose's result
        itself; return_code -1 on empty input."""
        res = CvoResultInfo()
        T0 = mat16(init_guess_T_s2t)
        Tout = np.zeros(16, np.float32)
        secs = C.c_double()
        iters = C.c_int32()
        ret = C.c_int32()
        a = Assoc()
        ap = C.byref(a) if is_returning_association else None
        if isinstance(source, DeviceCloud):
            st = lib().cvo_align_s2t_d(self._ctx, source.handle, target.handle, T0.ctypes.data, Tout.ctypes.data, ap,
                                       C.byref(secs), C.byref(iters), C.byref(ret))
        else:
            hs, ht = HeldView(source), HeldView(target)
            st = lib().cvo_align_s2t(self._ctx, hs.ref, ht.ref, T0.ctypes.data, Tout.ctypes.data, ap,
                                     C.byref(secs), C.byref(iters), C.byref(ret))
        if st == 2:
            res.return_code = -1
            return res
        _check(st, self._ctx)
        res.T_s2t, res.num_iters, res.registration_seconds, res.return_code = (Tout.reshape(4, 4), iters.value,
                                                                                secs.value, ret.value)
        if is_returning_association and a.row_ptr:
            res.association = _take_assoc(a)
        return res

    def get_nearest_init_pose(self, source, target, init_guess_ell: float, num_guess: int = 0):
        """get_nearest_init_pose (CvoGPU.cpp:367-398): the rotation of the
        60-rotation table with the largest approximate cos angle (first strict
        maximum above 0, else the identity) as a 4x4 source-to-target transform.
        Returns (T, seconds)."""
        T = np.zeros(16, np.float32)
        secs = C.c_double(0.0)
        if isinstance(source, DeviceCloud):
            st = lib().cvo_get_nearest_init_pose_d(self._ctx, source.handle, target.handle, C.c_float(init_guess_ell),
                                                   int(num_guess), T.ctypes.data, C.byref(secs))
        else:
            hs, ht = HeldView(source), HeldView(target)
            st = lib().cvo_get_nearest_init_pose(self._ctx, hs.ref, ht.ref, C.c_float(init_guess_ell), int(num_guess),
                                                 T.ctypes.data, C.byref(secs))
        _check(st, self._ctx)
        return T.reshape(4, 4), secs.value

    def align_multistart(self, source, target, T_starts, score_ell: float, mode: int = 0) -> dict:
        """Batched global registration: every start in T_starts (n,4,4, align
        init convention) is registered in one batched launch sequence (mode 0)
        and scored by the cos angle at score_ell; mode 1 scores the initial
        poses only and aligns from the best (get_nearest_init_pose,
        CvoGPU.cpp:367-398)."""
        own = []
        if not isinstance(source, DeviceCloud):
            source = self.upload(source)
            own.append(source)
        if not isinstance(target, DeviceCloud):
            target = self.upload(target)
            own.append(target)
        Ts = np.ascontiguousarray(np.asarray(T_starts, np.float32).reshape(-1, 16))
        n = Ts.shape[0]
        T_best = np.zeros(16, np.float32)
        per_T = np.zeros((n, 16), np.float32)
        per_cos = np.zeros(n, np.float64)
        per_it = np.zeros(n, np.int32)
        best = C.c_int32()
        secs = C.c_double()
        _check(lib().cvo_align_multistart(self._ctx, source.handle, target.handle, n, Ts.ctypes.data,
                                          C.c_float(score_ell), int(mode), T_best.ctypes.data, C.byref(best),
                                          per_T.ctypes.data, per_cos.ctypes.data, per_it.ctypes.data,
                                          C.byref(secs)), self._ctx)
        for c in own:
            c.free()
        return {"T_best": T_best.reshape(4, 4), "best_index": best.value, "per_start_T": per_T.reshape(n, 4, 4),
                "per_start_cos": per_cos, "per_start_iters": per_it, "seconds": secs.value}

    def align_multiframe(self, frames, poses, hold_const=None, edges=(), max_records: int = 1024,
                         inliers: bool = False) -> dict:
        """IRLS bundle adjustment of a window of frames (CvoGPU::align over
        frames, CvoGPU.cu:1614-1671, :1955-1999; CvoBatchIRLS::solve,
        IRLS.cpp:148-367).  frames: clouds (sensor frame) or DeviceClouds;
        poses: (n,3,4) or (n,4,4) world <- sensor; edges: (E,2) frame index
        pairs.  Returns the updated poses (n,4,4), the outer iteration count,
        the timings and, with inliers, each edge's Association of the last
        association pass (BinaryState::export_association)."""
        own = []
        dev = []
        for fr in frames:
            if isinstance(fr, DeviceCloud):
                dev.append(fr)
            else:
                d = self.upload(fr)
                own.append(d)
                dev.append(d)
        n = len(dev)
        P = np.ascontiguousarray(np.asarray(poses, np.float64)[:, :3, :4].reshape(n, 12))
        hc = None
        if hold_const is not None:
            hc = np.ascontiguousarray(np.asarray(hold_const, np.uint8).reshape(n))
        E = np.ascontiguousarray(np.asarray(edges, np.int32).reshape(-1, 2))
        handles = (C.c_void_p * max(n, 1))(*[d.handle.value for d in dev])
        secs = C.c_double()
        it = C.c_int32()
        per = np.zeros(max(1, max_records), np.float64)
        n_e = E.shape[0]
        per_edge = (Assoc * max(1, n_e))() if inliers else None
        try:
            _check(lib().cvo_align_multiframe(self._ctx, n, handles, P.ctypes.data,
                                              hc.ctypes.data if hc is not None else None, n_e,
                                              E.ctypes.data if E.size else None, per_edge, C.byref(secs),
                                              C.byref(it), per.ctypes.data, max_records), self._ctx)
        finally:
            for c in own:
                c.free()
        out = np.tile(np.eye(4), (n, 1, 1))
        out[:, :3, :4] = P.reshape(n, 3, 4)
        r = {"poses": out, "iters": it.value, "seconds": secs.value,
             "per_iter_seconds": per[:min(it.value + 1, max_records)]}
        if inliers:
            r["inliers"] = []
            for e in range(n_e):
                a = per_edge[e]
                if a.row_ptr:
                    r["inliers"].append(_take_assoc(a))
                else:
                    rows = dev[int(E[e, 0])].n
                    r["inliers"].append(Association(rows, dev[int(E[e, 1])].n, np.zeros(rows + 1, np.int64),
                                                    np.zeros(0, np.int32), np.zeros(0, np.float32)))
        return r

    def align_frames(self, frames, hold_const, edges) -> float:
        """CvoGPU::align(std::vector<CvoFrame::Ptr>&, hold_const, edges, seconds)
        (CvoGPU.hpp:156-164): frames carry `points` and a 12-double `pose_vec`
        that is updated in place; edges are pairs of frames.  Returns the
        seconds of the call."""
        idx = {id(f): i for i, f in enumerate(frames)}
        E = [(idx[id(a)], idx[id(b)]) for a, b in edges]
        poses = np.stack([np.asarray(f.pose_vec, np.float64).reshape(3, 4) for f in frames])
        r = self.align_multiframe([f.points for f in frames], poses, hold_const, E)
        for f, T in zip(frames, r["poses"]):
            np.copyto(f.pose_vec, T[:3, :4].reshape(12))
        return r["seconds"]

    def inner_product_gpu(self, source, target, init_guess_transform, ell: float) -> float:
        """inner_product_gpu (CvoGPU.cu:1734-1748)."""
        T0 = mat16(init_guess_transform)
        ip = C.c_double()
        if isinstance(source, DeviceCloud):
            st = lib().cvo_inner_product_d(self._ctx, source.handle, target.handle, T0.ctypes.data, C.c_float(ell),
                                           C.byref(ip))
        else:
            hs, ht = HeldView(source), HeldView(target)
            st = lib().cvo_inner_product(self._ctx, hs.ref, ht.ref, T0.ctypes.data, C.c_float(ell), C.byref(ip))
        if st == 2:
            return 0.0
        _check(st, self._ctx)
        return ip.value

    def function_angle(self, source, target, t2s_frame_transform, ell: float, is_approximate: bool = True,
                       is_gpu: bool = True) -> float:
        """function_angle (CvoGPU.cu:1768-1806).  is_gpu False gives the
        inner_product_cpu semantics (fixed lengthscale, no row cap, no gates
        but the radius); both run on the GPU."""
        T0 = mat16(t2s_frame_transform)
        hs, ht = HeldView(source), HeldView(target)
        out = C.c_double()
        _check(lib().cvo_function_angle_ex(self._ctx, hs.ref, ht.ref, T0.ctypes.data, C.c_float(ell),
                                           int(bool(is_approximate)), int(bool(is_gpu)), C.byref(out)), self._ctx)
        return out.value

    def inner_product_noalloc(self, source: "DeviceCloud", target: "DeviceCloud", init_guess_transform,
                              ell: float) -> float:
        """inner_product_gpu on device-resident clouds without any device
        allocation: the library's fixed per-GPU scratch array holds every
        intermediate (at most 131072 points per cloud; calls that overlap on
        one GPU race for it)."""
        T0 = mat16(init_guess_transform)
        ip = C.c_double()
        _check(lib().cvo_inner_product_d_noalloc(self._ctx, source.handle, target.handle, T0.ctypes.data,
                                                 C.c_float(ell), C.byref(ip)), self._ctx)
        return ip.value

    def inner_product_cpu(self, source, target, t2s_frame_transform, ell: float) -> float:
        """inner_product_cpu (CvoGPU.cpp:204-231), computed on the GPU with
        that function's semantics: the target moved by the inverse transform,
        every pair inside the radius at the fixed lengthscale, no row cap and
        no colour / semantic / geometric-type gate."""
        T0 = mat16(t2s_frame_transform)
        hs, ht = HeldView(source), HeldView(target)
        ip = C.c_double()
        _check(lib().cvo_inner_product_cpu(self._ctx, hs.ref, ht.ref, T0.ctypes.data, C.c_float(ell), C.byref(ip)),
               self._ctx)
        return ip.value

    def compute_association_gpu(self, source, target, T_target_frame_to_source_frame, lengthscale) -> Association:
        """compute_association_gpu (CvoGPU.cu:1836-1871); with a 3x3 matrix in
        place of the lengthscale, the non-isotropic overload (CvoGPU.cu:1873-1952)."""
        T0 = mat16(T_target_frame_to_source_frame)
        hs, ht = HeldView(source), HeldView(target)
        a = Assoc()
        if np.ndim(lengthscale) == 2:
            Km = np.ascontiguousarray(np.asarray(lengthscale, np.float32).reshape(9))
            _check(lib().cvo_compute_association_nonisotropic(self._ctx, hs.ref, ht.ref, T0.ctypes.data,
                                                              Km.ctypes.data, C.byref(a)), self._ctx)
        else:
            _check(lib().cvo_compute_association(self._ctx, hs.ref, ht.ref, T0.ctypes.data, C.c_float(lengthscale),
                                                 C.byref(a)), self._ctx)
        if not a.row_ptr:
            n = hs.view.n
            return Association(n, ht.view.n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
        return _take_assoc(a)

    def compute_association_knn(self, source, target, T_target_frame_to_source_frame, lengthscale,
                                K: int) -> Association:
        """The K-nearest association pass of align with is_using_kdtree
        (se_kernel_kdtree, CvoGPU.cu:1104-1154) at one pose, lengthscale and
        row cap K in [1, 512]."""
        T0 = mat16(T_target_frame_to_source_frame)
        hs, ht = HeldView(source), HeldView(target)
        a = Assoc()
        _check(lib().cvo_compute_association_knn(self._ctx, hs.ref, ht.ref, T0.ctypes.data, C.c_float(lengthscale),
                                                 int(K), C.byref(a)), self._ctx)
        if not a.row_ptr:
            n = hs.view.n
            return Association(n, ht.view.n, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
        return _take_assoc(a)

    # -- instrumentation ----------------------------------------------------
    def set_trace(self, cap: int) -> None:
        _check(lib().cvo_set_trace(self._ctx, int(cap)), self._ctx)

    def get_trace(self) -> List[IterRecord]:
        n = C.c_int32()
        _check(lib().cvo_get_trace(self._ctx, None, 0, C.byref(n)), self._ctx)
        buf = (IterRecord * max(1, n.value))()
        _check(lib().cvo_get_trace(self._ctx, buf, n.value, C.byref(n)), self._ctx)
        return [buf[i] for i in range(n.value)]

    def set_pose_arithmetic(self, mode: str) -> None:
        """"fp64" (default): FP64 pose and SE(3) exponential; "reference": the
        reference's float pose with its FP32 Exp_SEK3."""
        _check(lib().cvo_set_pose_arithmetic(self._ctx, {"fp64": 0, "reference": 1}[mode]), self._ctx)

    def set_profiling(self, on: bool) -> None:
        _check(lib().cvo_set_profiling(self._ctx, int(bool(on))), self._ctx)

    # -- workspace arena ------------------------------------------------------
    def workspace_bytes(self) -> int:
        n = C.c_uint64()
        _check(lib().cvo_workspace_bytes(self._ctx, C.byref(n)), self._ctx)
        return n.value

    def set_workspace_limit(self, nbytes: int) -> None:
        """Caps the context's device arena (0: no cap); a call that would
        exceed it raises CvoError (size beyond a documented limit)."""
        _check(lib().cvo_set_workspace_limit(self._ctx, C.c_uint64(int(nbytes))), self._ctx)

    def release_workspace(self) -> None:
        _check(lib().cvo_release_workspace(self._ctx), self._ctx)

    # -- association-kernel tuning ------------------------------------------
    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().cvo_get_tuning(self._ctx, C.byref(t)), self._ctx)
        return t.as_dict()

    def set_tuning(self, lanes_per_row: int = 16, xcd_remap: int = 1, cell_scale: float = 1.0,
                   lanes_small_ell: Optional[int] = None, split_ell: float = 0.0,
                   cell_scale_small_ell: Optional[float] = None) -> None:
        """Association launch shape: lanes_per_row (4, 8 or 16) and cell_scale at
        lengthscales >= split_ell, lanes_small_ell / cell_scale_small_ell below
        it (split_ell 0: one setting)."""
        t = _tuning(lanes_per_row, xcd_remap, cell_scale, lanes_small_ell, split_ell, cell_scale_small_ell)
        _check(lib().cvo_set_tuning(self._ctx, C.byref(t)), self._ctx)

    def apply_tuning(self, t: dict) -> None:
        """Sets a tuning dict as returned by autotune / get_tuning."""
        self.set_tuning(t["lanes_per_row"], t["xcd_remap"], t["cell_scale"], lanes_small_ell=t["lanes_small_ell"],
                        split_ell=t["split_ell"], cell_scale_small_ell=t["cell_scale_small_ell"])

    def autotune(self, source, target, init_guess_transform=None, cache: Optional[str] = "default") -> dict:
        """Times every association launch shape on this pair and keeps the
        fastest on the context.  cache: a JSON path, "default"
        (default_tuning_cache()) or None (measure, do not cache)."""
        T = mat16(np.eye(4) if init_guess_transform is None else init_guess_transform)
        path = default_tuning_cache() if cache == "default" else cache
        t = Tuning()
        cp = path.encode() if path else None
        if isinstance(source, DeviceCloud) and isinstance(target, DeviceCloud):
            _check(lib().cvo_autotune_d(self._ctx, source.handle, target.handle, T.ctypes.data, cp, C.byref(t)), self._ctx)
        else:
            hs, ht = HeldView(source), HeldView(target)
            _check(lib().cvo_autotune(self._ctx, hs.ref, ht.ref, T.ctypes.data, cp, C.byref(t)), self._ctx)
        return t.as_dict()

    def tuning_key(self, source: "DeviceCloud", target: "DeviceCloud") -> str:
        buf = C.create_string_buffer(512)
        _check(lib().cvo_tuning_key(self._ctx, source.handle, target.handle, buf, 512), self._ctx)
        return buf.value.decode()

    def kernel_stats(self) -> dict:
        s = KernelStats()
        _check(lib().cvo_get_kernel_stats(self._ctx, C.byref(s)), self._ctx)
        return {"ms": {k: s.ms[i] for i, k in enumerate(KSTAT_NAMES)},
                "launches": {k: s.launches[i] for i, k in enumerate(KSTAT_NAMES)},
                "bytes": {k: s.bytes[i] for i, k in enumerate(KSTAT_NAMES)},
                "pairs_stored": s.pairs_stored, "candidates": s.candidates}

    def reset_kernel_stats(self) -> None:
        _check(lib().cvo_reset_kernel_stats(self._ctx), self._ctx)


def version() -> str:
    return lib().cvo_version().decode()


def icosahedral_rotations() -> np.ndarray:
    """The 60 rotations of get_nearest_init_pose's table
    (gen_icosahedron_init_rots, CvoGPU.cpp:279-342) in the table's order,
    (60, 3, 3) float32, as the library holds them (no device needed)."""
    out = np.zeros(540, np.float32)
    _check(lib().cvo_icosahedral_rotations(out.ctypes.data))
    return out.reshape(60, 3, 3)
