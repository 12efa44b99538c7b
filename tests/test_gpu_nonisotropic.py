"""GPU non-isotropic association (cvo_compute_association_nonisotropic,
CvoGPU::compute_association_gpu with a 3x3 kernel, src/cvo/CvoGPU.cu:719-830,
:1873-1952) against the oracle's restatement on the same inputs.

Bars: identical (row, target) sets except pairs within 1e-6 relative of
sp_thres (the metric is formed in float on both sides); values within 1e-5 relative (the
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


def compare(cvo_gpu, oracle, p, src, tgt, T, Km, K):
    """GPU association against the oracle's; returns the common pair count."""
    p.nearest_neighbors_max = K
    cvo_gpu.write_params(p)
    a = cvo_gpu.compute_association_gpu(src, tgt, T, Km)
    counts, cols, vals = oracle.assoc_nonisotropic(p, src, tgt, T, Km, K)
    want = {(i, int(cols[i, s])): float(vals[i, s]) for i in range(src.n) for s in range(counts[i])}
    got = {}
    for i in range(a.rows):
        seg = a.col[a.row_ptr[i]:a.row_ptr[i + 1]]
        assert (np.diff(seg) > 0).all()
        assert len(seg) <= K
        for j, v in zip(seg, a.val[a.row_ptr[i]:a.row_ptr[i + 1]]):
            got[(i, int(j))] = float(v)
    diff = set(got) ^ set(want)
    band = max((abs(got.get(k, want.get(k, 0.0)) - p.sp_thres) / p.sp_thres for k in diff), default=0.0)
    say(f"  non-isotropic: {len(diff)} differing pairs, widest {band:.2e} of sp_thres")
    near = [k for k in diff if abs(got.get(k, want.get(k, 0.0)) - p.sp_thres) <= 1e-6 * p.sp_thres]
    assert len(diff) == len(near), (band, sorted(diff - set(near))[:10])
    common = set(got) & set(want)
    # an indefinite kernel's exp(-d2 / 2) overflows to inf on both sides
    rel = max((0.0 if got[k] == want[k] else abs(got[k] - want[k]) / want[k] for k in common), default=0.0)
    assert rel <= 1e-5, rel
    return len(common), int((counts == K).sum())


def box_rows(src, Km):
    """Rows whose metric takes the grid search (positive definite, with
    8u ||M||_F ||M^-1||_F <= 0.01); the others scan every target."""
    Kinv = np.linalg.inv(Km.astype(np.float64))
    x = src.xyz.astype(np.float64)
    z = x / np.linalg.norm(x, axis=1, keepdims=True)
    u = np.stack([np.ones(len(x)), np.ones(len(x)), -(z[:, 0] + z[:, 1]) / z[:, 2]], 1)
    U = np.stack([u, np.cross(z, u), z], axis=2)
    M = U @ Kinv[None] @ np.transpose(U, (0, 2, 1))
    if (np.linalg.eigvalsh(M) <= 0).any():
        return 0
    kap = np.linalg.norm(M, axis=(1, 2)) * np.linalg.norm(np.linalg.inv(M), axis=(1, 2))
    return int((8 * 2.0 ** -24 * kap <= 0.01).sum())


@pytest.mark.parametrize("kernel", [np.diag([0.02, 0.01, 0.04]),
                                    np.array([[0.03, 0.004, 0.0], [0.004, 0.015, 0.002], [0.0, 0.002, 0.05]])])
def test_nonisotropic_matches_oracle(cvo_gpu, oracle, kernel):
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    Km = kernel.astype(np.float32)
    say(f"non-isotropic association 5k/5k kernel diag {np.diag(Km)}")
    n, _ = compare(cvo_gpu, oracle, p, src, tgt, Tgt, Km, p.nearest_neighbors_max)
    assert n > 1000


# (kernel, cap, what the case exercises): rows over the cap in the grid
# search; a metric conditioned so that about half the rows take the grid
# search and half scan every target; an indefinite kernel (every row scans,
# every row over the cap)
CASES = [
    (np.diag([0.3, 0.15, 0.6]), 8, "capped rows, grid search"),
    (np.diag([3e-5, 0.3, 0.2]), 512, "grid search and full scans mixed"),
    (np.diag([0.02, -0.01, 0.04]), 8, "indefinite kernel, full scans"),
]


@pytest.mark.parametrize("kernel,K,what", CASES)
def test_nonisotropic_search_paths(cvo_gpu, oracle, kernel, K, what):
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    Km = kernel.astype(np.float32)
    nb = box_rows(src, Km)
    say(f"non-isotropic 5k/5k {what}: K={K}, {nb} of {src.n} rows grid-searched")
    n, capped = compare(cvo_gpu, oracle, p, src, tgt, Tgt, Km, K)
    assert n > 500
    if "capped" in what:
        assert capped > 1000 and nb > 4900
    if "mixed" in what:
        assert 1000 < nb < 4000
    if "indefinite" in what:
        assert nb == 0 and capped == src.n


def test_nonisotropic_street_semantics_capped(cvo_gpu, oracle):
    """19-class gate, a 5 deg / 1 m pose, 2k source rows against 20k targets,
    cap 5: rows over the cap keep their first 5 passing targets by index."""
    p = params_from_yaml("street_60k.yaml")
    src, tgt, Tgt = synth.street_pair(4, n=20000)
    src = src.subset(np.arange(0, src.n, 10))
    Km = np.diag([1.0, 0.2, 3.0]).astype(np.float32)
    say("non-isotropic street 2k x 20k, semantics, K=5")
    n, capped = compare(cvo_gpu, oracle, p, src, tgt, Tgt, Km, 5)
    assert n > 1000 and capped > 100


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
