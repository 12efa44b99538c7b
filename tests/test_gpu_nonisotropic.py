"""GPU non-isotropic association (cvo_compute_association_nonisotropic,
CvoGPU::compute_association_gpu with a 3x3 kernel, src/cvo/CvoGPU.cu:719-830,
:1873-1952) against the oracle's restatement on the same inputs.

Bars: identical (row, target) sets except pairs within 1e-4 relative of
sp_thres (the metric is formed in float); values within 1e-5 relative (the
same float formulas on both sides); target ids ascending within each row;
singular kernels rejected; empty input gives an empty association."""
import sys

import numpy as np
import pytest

from conftest import params_from_yaml
from rkhs_ba_amd import CvoError, synth
from rkhs_ba_amd.pointcloud import CvoPointCloud

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kernel", [np.diag([0.02, 0.01, 0.04]),
                                    np.array([[0.03, 0.004, 0.0], [0.004, 0.015, 0.002], [0.0, 0.002, 0.05]])])
def test_nonisotropic_matches_oracle(cvo_gpu, oracle, kernel):
    p = params_from_yaml("planes_spheres_5k.yaml")
    cvo_gpu.write_params(p)
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    Km = kernel.astype(np.float32)
    say(f"non-isotropic association 5k/5k kernel diag {np.diag(Km)}")
    a = cvo_gpu.compute_association_gpu(src, tgt, Tgt, Km)
    K = p.nearest_neighbors_max
    counts, cols, vals = oracle.assoc_nonisotropic(p, src, tgt, Tgt, Km, K)
    want = {(i, int(cols[i, s])): float(vals[i, s]) for i in range(src.n) for s in range(counts[i])}
    got = {}
    for i in range(a.rows):
        seg = a.col[a.row_ptr[i]:a.row_ptr[i + 1]]
        assert (np.diff(seg) > 0).all()
        for j, v in zip(seg, a.val[a.row_ptr[i]:a.row_ptr[i + 1]]):
            got[(i, int(j))] = float(v)
    diff = set(got) ^ set(want)
    near = [k for k in diff if abs(got.get(k, want.get(k, 0.0)) - p.sp_thres) <= 1e-4 * p.sp_thres]
    assert len(diff) == len(near), sorted(diff - set(near))[:10]
    common = set(got) & set(want)
    assert len(common) > 1000
    rel = max(abs(got[k] - want[k]) / want[k] for k in common)
    assert rel <= 1e-5, rel


def test_nonisotropic_edge_cases(cvo_gpu):
    p = params_from_yaml("planes_spheres_5k.yaml")
    cvo_gpu.write_params(p)
    xyz = np.array([[0, 0, 0], [1.0, 0.5, 0.0], [1.0, 0.2, 2.0]], np.float32)
    f = np.full((3, 5), 0.5, np.float32)
    s = CvoPointCloud.from_arrays(xyz, f)
    t = CvoPointCloud.from_arrays(xyz + 0.01, f)
    a = cvo_gpu.compute_association_gpu(s, t, np.eye(4), np.diag([0.05, 0.05, 0.05]))
    rows = np.diff(a.row_ptr)
    assert rows[0] == 0 and rows[1] == 0 and rows[2] >= 1
    with pytest.raises(CvoError):
        cvo_gpu.compute_association_gpu(s, t, np.eye(4), np.zeros((3, 3)))
    empty = CvoPointCloud.from_arrays(np.zeros((0, 3), np.float32), np.zeros((0, 5), np.float32))
    e = cvo_gpu.compute_association_gpu(s, empty, np.eye(4), np.eye(3) * 0.05)
    assert e.nnz == 0
