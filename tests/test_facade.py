"""The C++ facade (include/cvo/CvoGPU.hpp) compiles against the C ABI and links
the library; host-only checks run here, a GPU registration through it runs
under the gpu marker."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

YAML = os.path.join(GOLDEN, "yaml", "cvo_outdoor_params.yaml")


def facade_binary():
    from rkhs_ba_amd.build import FACADE_CHECK
    return FACADE_CHECK


def test_facade_compiles_and_runs_host_checks(tmp_path):
    exe = str(tmp_path / "facade_check")
    lib = os.path.join(ROOT, "rkhs_ba_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "facade_check.cpp"), "-L" + lib, "-lrkhs_ba_amd",
                    "-Wl,-rpath," + lib, "-o", exe], check=True)
    out = subprocess.run([exe, YAML], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ell_init 1 sp_thres 0.007 MAX_ITER 500" in out.stdout and "ok" in out.stdout


@pytest.mark.gpu
def test_facade_registers_on_the_gpu(tmp_path):
    # the same registration through the C++ facade and through the Python
    # binding of the C ABI gives the same transform bit for bit
    import numpy as np
    from rkhs_ba_amd import CvoGPU, synth
    exe = facade_binary()
    assert os.path.exists(exe), "build() compiles the facade check"
    dump = str(tmp_path / "facade.bin")
    print("facade check --gpu", flush=True)
    out = subprocess.run([exe, YAML, "--gpu", dump], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    raw = open(dump, "rb").read()
    n, ret, iters = np.frombuffer(raw[:12], np.int32)
    f = np.frombuffer(raw[12:], np.float32)
    ax, af = f[:3 * n].reshape(n, 3), f[3 * n:8 * n].reshape(n, 5)
    bx, bf = f[8 * n:11 * n].reshape(n, 3), f[11 * n:16 * n].reshape(n, 5)
    T_cpp = f[16 * n:16 * n + 16].reshape(4, 4)
    cvo = CvoGPU(YAML)
    r, T, _, it, _ = cvo.align(synth.Cloud(ax.copy(), af.copy()), synth.Cloud(bx.copy(), bf.copy()), np.eye(4))
    cvo.close()
    assert (r, it) == (ret, iters)
    np.testing.assert_array_equal(T, T_cpp)
