"""The reference's demo registration (README.md:62-67: demo_data/source.pcd ->
target.pcd with cvo_params/cvo_outdoor_params.yaml), color PCDs read by the
host CvoPointCloud.  With the reference's float pose arithmetic the GPU pose
equals the FP32-mirror oracle's, with the FP64 pose (default) the oracle's
restatement of that policy, both within 1e-4 rad / 1e-4 m and with the same
iteration count; the final association sets are equal."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rot_dist_rad, trans_err
from rkhs_ba_amd import CvoGPU, read_params_yaml
from rkhs_ba_amd.pointcloud import CvoPointCloud

pytestmark = pytest.mark.gpu


def test_demo_pcd_registration_matches_oracle(oracle):
    p = read_params_yaml(os.path.join(GOLDEN, "yaml", "cvo_outdoor_params.yaml"))
    src = CvoPointCloud.from_pcd(os.path.join(GOLDEN, "pcd", "source.pcd"))
    tgt = CvoPointCloud.from_pcd(os.path.join(GOLDEN, "pcd", "target.pcd"))
    cvo = CvoGPU(params=p)
    print("demo pcd: gpu align (reference pose arithmetic)", flush=True)
    cvo.set_pose_arithmetic("reference")
    ret_r, Tr, _, it_r, _ = cvo.align(src, tgt, np.eye(4))
    cvo.set_pose_arithmetic("fp64")
    To, it_o, ret_o, _ = oracle.align(p, src, tgt, np.eye(4))
    assert (ret_r, it_r) == (ret_o, it_o)
    assert rot_dist_rad(Tr, To) <= 1e-4 and trans_err(Tr, To) <= 1e-4, (rot_dist_rad(Tr, To), trans_err(Tr, To))
    print("demo pcd: gpu align (fp64 pose)", flush=True)
    ret, T, secs, iters, _ = cvo.align(src, tgt, np.eye(4))
    Tp, it_p, ret_p, _ = oracle.align(p, src, tgt, np.eye(4), precision=oracle.POLICY)
    assert (ret, iters) == (ret_p, it_p)
    assert rot_dist_rad(T, Tp) <= 1e-4 and trans_err(T, Tp) <= 1e-4, (rot_dist_rad(T, Tp), trans_err(T, Tp))
    # association at the converged pose and the final lengthscale of the run
    Tinv = np.linalg.inv(np.asarray(T, np.float64)).astype(np.float32)
    a_gpu = cvo.compute_association_gpu(src, tgt, Tinv, 0.1)
    counts, cols, vals, _ = oracle.assoc(p, src, tgt, Tinv, 0.1, p.nearest_neighbors_max)
    rows = np.repeat(np.arange(src.size()), counts)
    ref = set(zip(rows.tolist(), cols[cols >= 0].tolist()))
    got = set(zip(np.repeat(np.arange(src.size()), np.diff(a_gpu.row_ptr)).tolist(), a_gpu.col.tolist()))
    assert got == ref
    cvo.close()
