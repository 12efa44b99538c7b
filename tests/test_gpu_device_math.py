"""The device controller's math against published values, through the C ABI
(cvo_device_math_check): the closed-form cubic against numpy.roots
(tests/golden/cubic_roots.json: 47 cubics incl. complex pairs, repeated and
widely scaled roots), the FP64 SE(3) exponential and log norm against
scipy's expm / logm (tests/golden/se3_exp_log.json: 30 twists incl. near-zero
rotations), and the reference's float Exp_SEK3 with its small-angle quirk.
The same bars as the oracle's own tests of these functions
(tests/test_oracle_golden.py)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from rkhs_ba_amd import lib

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def device_math(coefs, xis, dts):
    nc, ne = len(coefs), len(xis)
    coef = np.ascontiguousarray(np.array(coefs, np.float64).reshape(nc, 4))
    xi = np.ascontiguousarray(np.array(xis, np.float64).reshape(ne, 6))
    dt = np.ascontiguousarray(np.array(dts, np.float64).reshape(ne))
    re, im = np.zeros((nc, 3)), np.zeros((nc, 3))
    T64, lg, T32 = np.zeros((ne, 3, 4)), np.zeros(ne), np.zeros((ne, 3, 4), np.float32)
    L = lib()
    L.cvo_device_math_check.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32] + [C.c_void_p] * 5
    st = L.cvo_device_math_check(nc, _p(coef), _p(re), _p(im), ne, _p(xi), _p(dt), _p(T64), _p(lg), _p(T32))
    assert st == 0, st
    return re + 1j * im, T64, lg, T32


def test_device_cubic_against_numpy_roots():
    cases = json.load(open(os.path.join(GOLDEN, "cubic_roots.json")))
    got, _, _, _ = device_math([c["coef"] for c in cases], [], [])
    for case, g in zip(cases, got):
        want = np.array(case["re"]) + 1j * np.array(case["im"])
        scale = max(1.0, np.max(np.abs(want)))
        for w in want:  # every numpy root matched by one device root
            assert np.min(np.abs(g - w)) <= 1e-6 * scale + 1e-9, (case["coef"], g, want)


def test_device_se3_exp_and_log_norm_against_scipy():
    cases = json.load(open(os.path.join(GOLDEN, "se3_exp_log.json")))
    _, T64, lg, T32 = device_math([], [c["xi"] for c in cases], [c["dt"] for c in cases])
    for case, T, n, f in zip(cases, T64, lg, T32):
        want = np.array(case["T"]).reshape(3, 4)
        assert np.allclose(T, want, atol=1e-12), case
        assert n == pytest.approx(case["log_norm"], rel=1e-7, abs=1e-12)
        if np.linalg.norm(case["xi"][:3]) >= 1e-6:  # the float form outside its small-angle branch
            assert np.allclose(f, want, atol=2e-5), case


def test_device_float_exp_small_angle_quirk():
    # below |w| = 1e-6 the reference's float form takes J = I: a pure
    # translation twist moves by v, not dt * v (LieGroup.cpp:252-254)
    _, T64, _, T32 = device_math([], [[0, 0, 0, 0.6, 0, 0.8]], [0.01])
    assert np.allclose(T32[0][:, 3], [0.6, 0.0, 0.8])
    assert np.allclose(T64[0][:, 3], [0.006, 0.0, 0.008], atol=1e-15)


def test_device_math_rejects_bad_arguments():
    L = lib()
    L.cvo_device_math_check.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32] + [C.c_void_p] * 5
    assert L.cvo_device_math_check(-1, None, None, None, 0, None, None, None, None, None) == 1
    assert L.cvo_device_math_check(2, None, None, None, 0, None, None, None, None, None) == 1
    assert L.cvo_device_math_check(0, None, None, None, 0, None, None, None, None, None) == 0
