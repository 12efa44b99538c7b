"""Host point cloud with the methods of cvo::CvoPointCloud
(include/UnifiedCvo/utils/CvoPointCloud.hpp:38-221, src/utils/CvoPointCloud.cpp)
without PCL or OpenCV: structure-of-arrays storage that the library uploads as
is (xyz, features, class distribution, geometric type), PCD and text I/O.

Feature and class widths follow the cloud (the reference fixes them at 5 and
20 at compile time and leaves the unused tail zero; the kernels compare only
the stored dimensions, which is the same distance).
"""
from __future__ import annotations

import struct
from typing import Optional

import numpy as np


class CvoPointCloud:
    """cvo::CvoPointCloud.  Attributes xyz (n,3), feat (n,F) or None, sem (n,C)
    or None, geo_type (n,2) are what CvoGPU uploads."""

    def __init__(self, feature_dimensions: int = 0, num_classes: int = 0, num_geometric_types: int = 2):
        self.feature_dimensions_ = int(feature_dimensions)
        self.num_classes_ = int(num_classes)
        self.num_geometric_types_ = int(num_geometric_types)
        self._xyz = np.zeros((0, 3), np.float32)
        self._feat = np.zeros((0, self.feature_dimensions_), np.float32)
        self._sem = np.zeros((0, self.num_classes_), np.float32)
        self._geo = np.zeros((0, self.num_geometric_types_), np.float32)

    # ---- construction ------------------------------------------------------
    @classmethod
    def from_arrays(cls, xyz, feat=None, sem=None, geo_type=None) -> "CvoPointCloud":
        xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
        c = cls(0 if feat is None else np.asarray(feat).shape[1], 0 if sem is None else np.asarray(sem).shape[1])
        c._xyz = xyz.copy()
        c._feat = np.zeros((len(xyz), 0), np.float32) if feat is None else np.asarray(feat, np.float32).copy()
        c._sem = np.zeros((len(xyz), 0), np.float32) if sem is None else np.asarray(sem, np.float32).copy()
        c._geo = (np.tile(np.array([0.0, 1.0], np.float32), (len(xyz), 1)) if geo_type is None
                  else np.asarray(geo_type, np.float32).copy())
        return c

    @classmethod
    def from_xyzrgb(cls, xyz, rgb) -> "CvoPointCloud":
        """CvoPointCloud(const pcl::PointCloud<pcl::PointXYZRGB>&)
        (src/utils/CvoPointCloud.cpp:576-602): features (b, g, r) / 255,
        geometric type (0, 1), no classes."""
        rgb = np.asarray(rgb).astype(np.uint32)
        r = ((rgb >> 16) & 0xFF).astype(np.float32)
        g = ((rgb >> 8) & 0xFF).astype(np.float32)
        b = (rgb & 0xFF).astype(np.float32)
        feat = np.stack([b, g, r], axis=1) / np.float32(255.0)
        return cls.from_arrays(xyz, feat, None, None)

    @classmethod
    def from_pcd(cls, path: str) -> "CvoPointCloud":
        """A PCD file (ascii or binary): x y z and, when present, packed rgb
        (as the XYZRGB constructor) or intensity (as a one-dimensional
        feature)."""
        fields, data = read_pcd(path)
        xyz = np.stack([data["x"], data["y"], data["z"]], axis=1).astype(np.float32)
        if "rgb" in data:
            return cls.from_xyzrgb(xyz, data["rgb"].view(np.uint32) if data["rgb"].dtype == np.float32
                                   else data["rgb"])
        if "intensity" in data:
            return cls.from_arrays(xyz, data["intensity"].astype(np.float32)[:, None], None, None)
        return cls.from_arrays(xyz, None, None, None)

    # ---- the reference's accessors ------------------------------------------
    def num_points(self) -> int:
        return int(self._xyz.shape[0])

    def size(self) -> int:
        return self.num_points()

    def num_classes(self) -> int:
        return self.num_classes_

    def num_features(self) -> int:
        return self.feature_dimensions_

    def feature_dimensions(self) -> int:
        return self.feature_dimensions_

    def num_geometric_types(self) -> int:
        return self.num_geometric_types_

    def positions(self) -> np.ndarray:
        return self._xyz.copy()

    def at(self, i: int) -> np.ndarray:
        return self._xyz[i].copy()

    xyz_at = at

    def feature_at(self, i: int) -> np.ndarray:
        return self._feat[i].copy()

    def label_at(self, i: int) -> np.ndarray:
        return self._sem[i].copy()

    def geometry_type_at(self, i: int) -> np.ndarray:
        return self._geo[i].copy()

    # ---- editing ------------------------------------------------------------
    def reserve(self, num_points: int, feature_dims: int, num_classes: int) -> None:
        self.feature_dimensions_ = int(feature_dims)
        self.num_classes_ = int(num_classes)
        self._xyz = np.zeros((num_points, 3), np.float32)
        self._feat = np.zeros((num_points, feature_dims), np.float32)
        self._sem = np.zeros((num_points, num_classes), np.float32)
        self._geo = np.zeros((num_points, self.num_geometric_types_), np.float32)

    def add_point(self, index: int, xyz, feature, label, geometric_type) -> int:
        """add_point (src/utils/CvoPointCloud.cpp:1579-1620) into a reserved slot."""
        if index < 0 or index >= self.num_points():
            return -1
        self._xyz[index] = xyz
        if self.feature_dimensions_:
            self._feat[index] = np.asarray(feature, np.float32)[:self.feature_dimensions_]
        if self.num_classes_:
            self._sem[index] = np.asarray(label, np.float32)[:self.num_classes_]
        self._geo[index] = np.asarray(geometric_type, np.float32)[:self.num_geometric_types_]
        return 0

    def push_back(self, xyz, feature=None, label=None, geometric_type=(0.0, 1.0)) -> None:
        n = self.num_points()
        self._xyz = np.vstack([self._xyz, np.asarray(xyz, np.float32).reshape(1, 3)])
        self._feat = np.vstack([self._feat, np.zeros((1, self.feature_dimensions_), np.float32)])
        self._sem = np.vstack([self._sem, np.zeros((1, self.num_classes_), np.float32)])
        self._geo = np.vstack([self._geo, np.zeros((1, self.num_geometric_types_), np.float32)])
        self.add_point(n, xyz, feature if feature is not None else np.zeros(self.feature_dimensions_),
                       label if label is not None else np.zeros(self.num_classes_), geometric_type)

    def erase(self, index: int) -> None:
        keep = np.ones(self.num_points(), bool)
        keep[index] = False
        self._xyz, self._feat, self._sem, self._geo = (self._xyz[keep], self._feat[keep], self._sem[keep],
                                                        self._geo[keep])

    def clear(self) -> None:
        self.reserve(0, self.feature_dimensions_, self.num_classes_)

    def __add__(self, b: "CvoPointCloud") -> "CvoPointCloud":
        if (self.feature_dimensions_, self.num_classes_) != (b.feature_dimensions_, b.num_classes_):
            raise ValueError("point clouds differ in feature or class dimension")
        c = CvoPointCloud(self.feature_dimensions_, self.num_classes_, self.num_geometric_types_)
        c._xyz = np.vstack([self._xyz, b._xyz])
        c._feat = np.vstack([self._feat, b._feat])
        c._sem = np.vstack([self._sem, b._sem])
        c._geo = np.vstack([self._geo, b._geo])
        return c

    @staticmethod
    def transform(pose, cloud: "CvoPointCloud") -> "CvoPointCloud":
        """CvoPointCloud::transform (src/utils/CvoPointCloud.cpp:1529-1558):
        positions through a 4x4 pose, attributes kept."""
        T = np.asarray(pose, np.float64)
        out = CvoPointCloud.from_arrays(cloud._xyz, cloud.feat, cloud.sem, cloud._geo)
        out._xyz = (cloud._xyz.astype(np.float64) @ T[:3, :3].T + T[:3, 3]).astype(np.float32)
        return out

    # ---- what the library uploads -----------------------------------------
    @property
    def n(self) -> int:
        return self.num_points()

    @property
    def xyz(self) -> np.ndarray:
        return self._xyz

    @property
    def feat(self) -> Optional[np.ndarray]:
        return self._feat if self.feature_dimensions_ else None

    @property
    def sem(self) -> Optional[np.ndarray]:
        return self._sem if self.num_classes_ else None

    @property
    def geo_type(self) -> np.ndarray:
        return self._geo

    # ---- text I/O ------------------------------------------------------------
    def read_cvo_pointcloud_from_file(self, filename: str) -> int:
        """read_cvo_pointcloud_from_file (src/utils/CvoPointCloud.cpp:1230-1260):
        header "num_points num_features num_classes", then per point
        u v idepth, features, x y z, class distribution.  Returns 0, or -1 if
        the file cannot be read."""
        try:
            tok = open(filename).read().split()
        except OSError:
            return -1
        n, F, Cn = int(tok[0]), int(tok[1]), int(tok[2])
        vals = np.array(tok[3:3 + n * (3 + F + 3 + Cn)], np.float64).reshape(n, 3 + F + 3 + Cn)
        self.reserve(n, F, Cn)
        self._feat[:] = vals[:, 3:3 + F]
        self._xyz[:] = vals[:, 3 + F:6 + F]
        self._sem[:] = vals[:, 6 + F:]
        self._geo[:] = (0.0, 1.0)
        return 0

    def write_cvo_pointcloud_file(self, filename: str) -> None:
        """The layout read_cvo_pointcloud_from_file reads (u, v, idepth written
        as 0)."""
        with open(filename, "w") as f:
            f.write(f"{self.num_points()} {self.feature_dimensions_} {self.num_classes_}\n")
            for i in range(self.num_points()):
                row = [0.0, 0.0, 0.0, *self._feat[i], *self._xyz[i], *self._sem[i]]
                f.write(" ".join(repr(float(np.float32(v))) for v in row) + "\n")

    def write_to_txt(self, name: str) -> None:
        """write_to_txt (src/utils/CvoPointCloud.cpp:1488-1510): header
        "num_points num_classes", then per point "x y z" on one line and the
        features and classes on the next."""
        with open(name, "w") as f:
            f.write(f"{self.num_points()} {self.num_classes_}\n")
            for i in range(self.num_points()):
                f.write(" ".join(f"{v:g}" for v in self._xyz[i]) + "\n")
                vals = list(self._feat[i]) + (list(self._sem[i]) if self.num_classes_ else [])
                f.write("".join(f"{v:g} " for v in vals) + "\n")

    def write_to_pcd(self, name: str) -> None:
        """x y z as an ascii PCD (write_to_pcd, src/utils/CvoPointCloud.cpp)."""
        write_pcd(name, {"x": self._xyz[:, 0], "y": self._xyz[:, 1], "z": self._xyz[:, 2]})


# ---------------------------------------------------------------------------
# PCD
# ---------------------------------------------------------------------------

_PCD_TYPES = {("F", 4): np.float32, ("F", 8): np.float64, ("U", 1): np.uint8, ("U", 2): np.uint16,
              ("U", 4): np.uint32, ("I", 1): np.int8, ("I", 2): np.int16, ("I", 4): np.int32}


def read_pcd(path: str):
    """PCD v0.7, DATA ascii or binary (not binary_compressed).  Returns the
    field names and a dict of column arrays (COUNT 1 fields)."""
    with open(path, "rb") as f:
        header = {}
        while True:
            line = f.readline()
            if not line:
                raise ValueError("PCD header without DATA")
            s = line.decode("ascii", "replace").strip()
            if not s or s.startswith("#"):
                continue
            key, *vals = s.split()
            header[key.upper()] = vals
            if key.upper() == "DATA":
                break
        body = f.read()
    fields = header["FIELDS"]
    sizes = [int(v) for v in header["SIZE"]]
    types = header["TYPE"]
    counts = [int(v) for v in header.get("COUNT", ["1"] * len(fields))]
    n = int(header["POINTS"][0])
    mode = header["DATA"][0].lower()
    dt = np.dtype([(fld if c == 1 else f"{fld}", _PCD_TYPES[(t, sz)], (c,) if c > 1 else ())
                   for fld, sz, t, c in zip(fields, sizes, types, counts)])
    if mode == "ascii":
        rows = np.array(body.decode("ascii").split(), np.float64).reshape(n, sum(counts))
        out, col = {}, 0
        for fld, sz, t, c in zip(fields, sizes, types, counts):
            v = rows[:, col:col + c]
            col += c
            arr = v[:, 0] if c == 1 else v
            out[fld] = arr.astype(_PCD_TYPES[(t, sz)]) if t != "F" else arr.astype(np.float32 if sz == 4 else np.float64)
        return fields, out
    if mode == "binary":
        rec = np.frombuffer(body[:n * dt.itemsize], dtype=dt, count=n)
        return fields, {fld: np.array(rec[fld]) for fld in fields}
    raise ValueError(f"PCD DATA {mode} is not supported")


def write_pcd(path: str, columns: dict) -> None:
    names = list(columns)
    n = len(columns[names[0]])
    with open(path, "w") as f:
        f.write("# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\n")
        f.write("FIELDS " + " ".join(names) + "\n")
        f.write("SIZE " + " ".join("4" for _ in names) + "\nTYPE " + " ".join("F" for _ in names) + "\n")
        f.write("COUNT " + " ".join("1" for _ in names) + f"\nWIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\n")
        f.write(f"POINTS {n}\nDATA ascii\n")
        for i in range(n):
            f.write(" ".join(repr(float(np.float32(columns[k][i]))) for k in names) + "\n")
