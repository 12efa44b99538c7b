"""Host point cloud (cvo::CvoPointCloud methods) and its PCD / text I/O on the
reference's demo data (tests/golden/pcd: demo_data/source.pcd, target.pcd,
bunny.pcd of the reference repository)."""
import os
import struct

import numpy as np

from conftest import GOLDEN
from rkhs_ba_amd.pointcloud import CvoPointCloud, read_pcd

PCD = os.path.join(GOLDEN, "pcd")


def first_data_line(path):
    with open(path, "rb") as f:
        for line in f:
            if line.startswith(b"DATA"):
                return f.readline().decode().split()


def test_ascii_pcd_xyzrgb_matches_the_xyzrgb_constructor():
    for name, n in (("source.pcd", 523), ("target.pcd", 1080)):
        c = CvoPointCloud.from_pcd(os.path.join(PCD, name))
        assert c.size() == n and c.num_features() == 3 and c.num_classes() == 0
        x, y, z, rgb = first_data_line(os.path.join(PCD, name))
        np.testing.assert_array_equal(c.at(0), np.array([x, y, z], np.float32))
        v = int(rgb)
        # features (b, g, r) / 255 (CvoPointCloud.cpp:592-594)
        np.testing.assert_array_equal(c.feature_at(0), np.array([v & 255, (v >> 8) & 255, (v >> 16) & 255],
                                                                np.float32) / np.float32(255.0))
        np.testing.assert_array_equal(c.geometry_type_at(0), [0.0, 1.0])
        assert (c.feat >= 0).all() and (c.feat <= 1).all()


def test_binary_pcd():
    path = os.path.join(PCD, "bunny.pcd")
    fields, cols = read_pcd(path)
    assert fields == ["x", "y", "z", "normal_x", "normal_y", "normal_z"]
    raw = open(path, "rb").read()
    off = raw.index(b"DATA binary\n") + len(b"DATA binary\n")
    first = struct.unpack("<6f", raw[off:off + 24])
    last = struct.unpack("<6f", raw[off + 24 * 14805:off + 24 * 14806])
    c = CvoPointCloud.from_pcd(path)
    assert c.size() == 14806 and c.feat is None
    np.testing.assert_array_equal(c.at(0), np.array(first[:3], np.float32))
    np.testing.assert_array_equal(c.at(14805), np.array(last[:3], np.float32))
    np.testing.assert_array_equal(cols["normal_z"][0], np.float32(first[5]))


def test_text_round_trip_and_write_to_txt(tmp_path):
    rng = np.random.default_rng(0)
    c = CvoPointCloud.from_arrays(rng.normal(size=(50, 3)), rng.random((50, 5)), rng.random((50, 19)))
    p = str(tmp_path / "c.txt")
    c.write_cvo_pointcloud_file(p)
    d = CvoPointCloud(0, 0)
    assert d.read_cvo_pointcloud_from_file(p) == 0
    np.testing.assert_array_equal(d.xyz, c.xyz)
    np.testing.assert_array_equal(d.feat, c.feat)
    np.testing.assert_array_equal(d.sem, c.sem)
    assert d.read_cvo_pointcloud_from_file(str(tmp_path / "missing.txt")) == -1
    q = str(tmp_path / "w.txt")
    c.write_to_txt(q)
    lines = open(q).read().splitlines()
    assert lines[0] == "50 19" and len(lines) == 1 + 2 * 50
    assert len(lines[1].split()) == 3 and len(lines[2].split()) == 5 + 19


def test_editing_and_transform():
    c = CvoPointCloud(5, 0)
    c.reserve(3, 5, 0)
    for i in range(3):
        assert c.add_point(i, [i, 0, 0], np.full(5, i / 10), [], [0, 1]) == 0
    assert c.add_point(3, [0, 0, 0], np.zeros(5), [], [0, 1]) == -1
    c.push_back([9, 9, 9], np.ones(5))
    assert c.size() == 4 and np.allclose(c.at(3), 9)
    c.erase(1)
    assert c.size() == 3 and np.allclose(c.at(1), [2, 0, 0])
    both = c + c
    assert both.size() == 6
    T = np.eye(4)
    T[:3, 3] = [1, 2, 3]
    moved = CvoPointCloud.transform(T, c)
    np.testing.assert_allclose(moved.xyz - c.xyz, np.tile([1, 2, 3], (3, 1)))
    np.testing.assert_array_equal(moved.feat, c.feat)
