"""The library's 60-rotation table of get_nearest_init_pose (host code, no
device): the reference's table (gen_icosahedron_init_rots, CvoGPU.cpp:279-342,
its float literals committed as a fixture) entry by entry and in order, a
group (closed under products, orthonormal, det +1)."""
import json
import os

import numpy as np

from conftest import GOLDEN
from rkhs_ba_amd import icosahedral_rotations


def test_table_matches_the_reference_order():
    ref = np.array(json.load(open(os.path.join(GOLDEN, "icosahedral_rotations.json"))), np.float64).reshape(60, 3, 3)
    R = icosahedral_rotations().astype(np.float64)
    assert np.abs(R - ref).max() < 1e-6
    # every entry is nearest to the fixture entry at the same position
    d = np.abs(R[:, None] - ref[None]).max(axis=(2, 3))
    assert (d.argmin(axis=1) == np.arange(60)).all()


def test_table_is_the_icosahedral_group():
    R = icosahedral_rotations().astype(np.float64)
    for M in R:
        assert np.abs(M @ M.T - np.eye(3)).max() < 1e-6
        assert abs(np.linalg.det(M) - 1) < 1e-6
    # closed: every product is an element
    flat = R.reshape(60, 9)
    for a in range(60):
        P = (R[a] @ R).reshape(60, 9)
        d = np.abs(P[:, None, :] - flat[None]).max(axis=2)
        assert (d.min(axis=1) < 1e-5).all()
    assert np.abs(R[29] - np.eye(3)).max() < 1e-7  # the identity sits where the reference has it
