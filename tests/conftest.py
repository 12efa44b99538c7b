import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CONFIGS = os.path.join(ROOT, "rkhs_ba_amd", "configs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def params_from_yaml(name):
    from rkhs_ba_amd import read_params_yaml
    return read_params_yaml(os.path.join(CONFIGS, name))


@pytest.fixture(scope="session")
def cvo_gpu():
    """One library context for the GPU tests (fails loudly without the .so or a GPU)."""
    from rkhs_ba_amd import CvoGPU
    c = CvoGPU(os.path.join(CONFIGS, "planes_spheres_5k.yaml"))
    yield c
    c.close()


def rot_err_deg(A, B):
    E = np.asarray(A, np.float64)[:3, :3].T @ np.asarray(B, np.float64)[:3, :3]
    return float(np.degrees(np.arccos(np.clip((np.trace(E) - 1) / 2, -1, 1))))


def trans_err(A, B):
    return float(np.linalg.norm(np.asarray(A, np.float64)[:3, 3] - np.asarray(B, np.float64)[:3, 3]))


def rot_dist_rad(A, B):
    """Angle between two rotations from the Frobenius distance of the matrices
    (2 asin(|RA - RB|_F / (2 sqrt 2))): linear near zero, so it resolves small
    angles between float32 matrices, unlike the arccos of the trace."""
    d = np.linalg.norm(np.asarray(A, np.float64)[:3, :3] - np.asarray(B, np.float64)[:3, :3])
    return float(2.0 * np.arcsin(min(1.0, d / (2.0 * np.sqrt(2.0)))))
