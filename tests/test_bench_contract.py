"""bench.py's host-side contract (CPU): more than one GPU is refused with one
JSON line and a non-zero exit, and its pose-difference helper is exact on
constructed poses."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 4 and "one GPU" in line["error"]


def test_rot_trans_diff_on_constructed_poses():
    import bench
    from rkhs_ba_amd import synth
    A = synth.rigid(synth.rot_axis_angle(np.array([0.3, -0.5, 0.8]), 0.7), np.array([1.0, -2.0, 0.5]))
    # B = A . [R(axis, 0.01) | t] with |t| = 0.25: the difference is that relative motion
    D = synth.rigid(synth.rot_axis_angle(np.array([1.0, 2.0, -1.0]), 0.01), np.array([0.0, 0.15, 0.2]))
    dr, dt = bench.rot_trans_diff(A, A @ D)
    assert abs(dr - 0.01) < 1e-12 and abs(dt - 0.25) < 1e-12
    dr, dt = bench.rot_trans_diff(A, A)
    assert dr < 1e-7 and dt < 1e-12  # arccos near 1 is coarse: a 1e-15 residue reads as ~1e-8 rad
