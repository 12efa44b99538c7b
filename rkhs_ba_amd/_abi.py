"""ctypes mirror of the structs in include/cvo_c.h (layout only, no behaviour)."""
from __future__ import annotations

import ctypes as C

import numpy as np

_PARAM_FIELDS = [
    ("ell_init_first_frame", C.c_float),
    ("ell_init", C.c_float),
    ("ell_min", C.c_float),
    ("min_ell_iter_limit", C.c_int32),
    ("ell_max", C.c_float),
    ("dl", C.c_double),
    ("dl_step", C.c_double),
    ("sigma", C.c_float),
    ("sp_thres", C.c_float),
    ("c", C.c_float),
    ("d", C.c_float),
    ("c_ell", C.c_float),
    ("c_sigma", C.c_float),
    ("s_ell", C.c_float),
    ("s_sigma", C.c_float),
    ("MAX_ITER", C.c_int32),
    ("eps", C.c_float),
    ("eps_2", C.c_float),
    ("min_step", C.c_float),
    ("max_step", C.c_float),
    ("step", C.c_float),
    ("nearest_neighbors_max", C.c_int32),
    ("ell_decay_rate", C.c_float),
    ("ell_decay_rate_first_frame", C.c_float),
    ("ell_decay_start", C.c_int32),
    ("ell_decay_start_first_frame", C.c_int32),
    ("indicator_window_size", C.c_int32),
    ("indicator_stable_threshold", C.c_float),
    ("is_pcl_visualization_on", C.c_int32),
    ("is_using_least_square", C.c_int32),
    ("is_ell_adaptive", C.c_int32),
    ("is_full_ip_matrix", C.c_int32),
    ("is_using_geometry", C.c_int32),
    ("is_using_intensity", C.c_int32),
    ("is_using_semantics", C.c_int32),
    ("is_using_range_ell", C.c_int32),
    ("is_using_kdtree", C.c_int32),
    ("is_exporting_association", C.c_int32),
    ("is_using_geometric_type", C.c_int32),
    ("is_global_angle_registration", C.c_int32),
    ("multiframe_using_cpu", C.c_int32),
    ("multiframe_max_iters", C.c_int32),
    ("multiframe_ell_init", C.c_float),
    ("multiframe_ell_min", C.c_float),
    ("multiframe_ell_max", C.c_float),
    ("multiframe_iter_per_ell", C.c_int32),
    ("multiframe_ell_decay_rate", C.c_float),
    ("multiframe_iterations_per_ell", C.c_int32),
    ("multiframe_iterations_per_solve", C.c_int32),
    ("multiframe_expected_points", C.c_int32),
    ("multiframe_downsample_voxel_size", C.c_float),
    ("multiframe_num_neighbors", C.c_int32),
    ("multiframe_least_squares_num_threads", C.c_int32),
    ("multiframe_min_nonzeros", C.c_int32),
    ("multiframe_is_optimizing_ell", C.c_int32),
    ("multiframe_is_sorting_inner_product", C.c_int32),
    ("multiframe_is_free_memory_each_iter", C.c_int32),
]


class CvoParams(C.Structure):
    """The cvo_params struct (reference CvoParams, CvoParams.hpp:12-79)."""
    _fields_ = _PARAM_FIELDS

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in _PARAM_FIELDS}

    def copy(self) -> "CvoParams":
        p = CvoParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(CvoParams))
        return p


PARAM_NAMES = [k for k, _ in _PARAM_FIELDS]


class CloudView(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("feat_dim", C.c_int32),
        ("num_classes", C.c_int32),
        ("xyz", C.POINTER(C.c_float)),
        ("feat", C.POINTER(C.c_float)),
        ("sem", C.POINTER(C.c_float)),
        ("geo_type", C.POINTER(C.c_float)),
    ]


class Assoc(C.Structure):
    _fields_ = [
        ("rows", C.c_int32),
        ("cols", C.c_int32),
        ("nnz", C.c_int64),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)),
        ("val", C.POINTER(C.c_float)),
    ]


class IterRecord(C.Structure):
    _fields_ = [
        ("ell", C.c_float),
        ("step", C.c_float),
        ("ip", C.c_float),
        ("nnz", C.c_int32),
        ("max_row", C.c_int32),
        ("num_neighbors", C.c_int32),
        ("dist", C.c_float),
        ("flow_norm", C.c_float),
    ]


KSTAT_COUNT = 8


class KernelStats(C.Structure):
    _fields_ = [
        ("ms", C.c_double * KSTAT_COUNT),
        ("launches", C.c_int64 * KSTAT_COUNT),
        ("bytes", C.c_double * KSTAT_COUNT),
        ("pairs_stored", C.c_double),
        ("candidates", C.c_double),
    ]


KSTAT_NAMES = ["assoc", "overflow", "step", "grid", "ba_pair", "ba_solve", "mlp", "aniso"]


class Tuning(C.Structure):
    """cvo_tuning: launch shape of the association kernel."""
    _fields_ = [
        ("lanes_per_row", C.c_int32),
        ("lanes_small_ell", C.c_int32),
        ("split_ell", C.c_float),
        ("xcd_remap", C.c_int32),
        ("cell_scale", C.c_float),
        ("cell_scale_small_ell", C.c_float),
        ("source", C.c_int32),
        ("score_us", C.c_float),
        ("n_tried", C.c_int32),
    ]

    SOURCES = {0: "default", 1: "set", 2: "measured", 3: "cached"}

    def as_dict(self) -> dict:
        return {"lanes_per_row": self.lanes_per_row, "lanes_small_ell": self.lanes_small_ell,
                "split_ell": round(float(self.split_ell), 6), "xcd_remap": self.xcd_remap,
                "cell_scale": round(float(self.cell_scale), 4),
                "cell_scale_small_ell": round(float(self.cell_scale_small_ell), 4),
                "source": self.SOURCES.get(self.source, self.source),
                "score_us": round(float(self.score_us), 3), "n_tried": self.n_tried}


def fptr(a):
    if a is None:
        return C.POINTER(C.c_float)()
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


class HeldView:
    """A CloudView plus the contiguous float32 arrays it points into."""

    def __init__(self, cloud):
        self.xyz = np.ascontiguousarray(cloud.xyz, dtype=np.float32)
        self.feat = None if cloud.feat is None else np.ascontiguousarray(cloud.feat, dtype=np.float32)
        self.sem = None if cloud.sem is None else np.ascontiguousarray(cloud.sem, dtype=np.float32)
        self.geo = None if getattr(cloud, "geo_type", None) is None else np.ascontiguousarray(cloud.geo_type, dtype=np.float32)
        self.view = CloudView(
            self.xyz.shape[0],
            0 if self.feat is None else self.feat.shape[1],
            0 if self.sem is None else self.sem.shape[1],
            fptr(self.xyz), fptr(self.feat), fptr(self.sem), fptr(self.geo))

    @property
    def ref(self):
        return C.byref(self.view)


def mat16(T) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(T, dtype=np.float32).reshape(16))
