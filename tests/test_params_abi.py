"""CPU tests of the parameter reader and the C ABI surface (no GPU calls)."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

from conftest import CONFIGS, GOLDEN, ROOT

REFERENCE_DEFAULTS = dict(
    # CvoParams::CvoParams() (include/UnifiedCvo/cvo/CvoParams.hpp:80-136)
    ell_init_first_frame=0.5, ell_init=0.5, ell_min=0.05, min_ell_iter_limit=1, ell_max=1.2, dl=0.0, dl_step=0.3,
    sigma=0.1, sp_thres=0.0006, c=7.0, d=7.0, c_ell=0.15, c_sigma=0.6, s_ell=0.1, s_sigma=0.8, MAX_ITER=10000,
    min_step=2e-5, eps=0.00005, eps_2=0.000012, ell_decay_rate=0.9, ell_decay_rate_first_frame=0.99,
    ell_decay_start=30, ell_decay_start_first_frame=300, indicator_window_size=15, indicator_stable_threshold=0.2,
    is_pcl_visualization_on=0, is_using_least_square=0, is_ell_adaptive=0, is_full_ip_matrix=0, is_using_geometry=1,
    is_using_intensity=0, is_using_semantics=0, is_using_range_ell=0, is_using_kdtree=0, is_using_geometric_type=0,
    is_global_angle_registration=0, is_exporting_association=0, multiframe_using_cpu=1, multiframe_max_iters=200,
    nearest_neighbors_max=512, multiframe_ell_init=0.15, multiframe_ell_min=0.05, multiframe_ell_max=5.0,
    multiframe_iter_per_ell=10, multiframe_ell_decay_rate=0.7, multiframe_iterations_per_ell=50,
    multiframe_iterations_per_solve=8, multiframe_downsample_voxel_size=0.5, multiframe_expected_points=1000,
    multiframe_num_neighbors=128, multiframe_min_nonzeros=300, multiframe_least_squares_num_threads=24,
    multiframe_is_optimizing_ell=0, multiframe_is_sorting_inner_product=0, multiframe_is_free_memory_each_iter=0)

# keys the reference reader never reads (CvoParams.hpp:124 vs :298-299)
NOT_READ = {"multiframe_iter_per_ell", "step", "is_ell_adaptive"}


def _close(a, b):
    return np.float32(a) == np.float32(b) if isinstance(b, float) else a == b


def test_defaults_match_reference():
    from rkhs_ba_amd import default_params
    p = default_params()
    for k, v in REFERENCE_DEFAULTS.items():
        assert _close(getattr(p, k), v), k
    # the reference leaves max_step uninitialised; the documented default here
    assert np.float32(p.max_step) == np.float32(0.01)


@pytest.mark.parametrize("fn", sorted(os.listdir(os.path.join(GOLDEN, "yaml"))))
def test_reference_yaml_files(fn):
    """Every reference cvo_params YAML, read on top of the defaults, equals the
    PyYAML parse of the same file (golden) for keys present, defaults otherwise."""
    from rkhs_ba_amd import default_params, read_params_yaml, PARAM_NAMES
    gold = json.load(open(os.path.join(GOLDEN, "params_yaml.json")))[fn]
    p = read_params_yaml(os.path.join(GOLDEN, "yaml", fn))
    d = default_params()
    for k in PARAM_NAMES:
        got = getattr(p, k)
        if k in gold and k not in NOT_READ:
            want = gold[k]
            if isinstance(got, float):
                assert np.float32(got) == np.float32(want) or got == pytest.approx(want, rel=1e-12), (k, got, want)
            else:
                assert got == int(want), (k, got, want)
        else:
            assert got == getattr(d, k), k


def test_yaml_missing_file_and_bad_value(tmp_path):
    from rkhs_ba_amd import CvoError, read_params_yaml
    with pytest.raises(CvoError) as e:
        read_params_yaml(str(tmp_path / "nope.yaml"))
    assert e.value.status == 3
    f = tmp_path / "bad.yaml"
    f.write_text("%YAML:1.0\n---\nell_init: abc\nsigma: 0.2\nunknown_key: 3\n")
    with pytest.raises(CvoError) as e:
        read_params_yaml(str(f))
    assert e.value.status == 4


def test_yaml_unknown_keys_and_second_document(tmp_path):
    from rkhs_ba_amd import read_params_yaml
    f = tmp_path / "x.yaml"
    f.write_text("%YAML:1.0\n---\nsigma: 0.2 # comment\nfoo: bar\nMAX_ITER: 7\n---\nsigma: 0.9\n")
    p = read_params_yaml(str(f))
    assert np.float32(p.sigma) == np.float32(0.2) and p.MAX_ITER == 7


def test_workload_configs_parse():
    from rkhs_ba_amd import read_params_yaml
    for fn in os.listdir(CONFIGS):
        p = read_params_yaml(os.path.join(CONFIGS, fn))
        assert p.sp_thres > 0


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for m in re.finditer(r"\b(cvo_[a-z0-9_]+)\s*\(", txt):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from rkhs_ba_amd import lib
    L = lib()
    declared = _declared_symbols()
    assert len(declared) > 20
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing


def test_status_strings_and_version():
    from rkhs_ba_amd import lib
    L = lib()
    for s in range(10):
        assert L.cvo_status_string(s)
    assert b"gfx950" in L.cvo_version()


def test_device_math_check_validates_before_the_device():
    """Bad sizes and missing buffers fail with CVO_ERR_INVALID_ARG, and an empty
    request succeeds, without a device call (so on a host without a GPU)."""
    from rkhs_ba_amd import lib
    L = lib()
    L.cvo_device_math_check.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32] + [C.c_void_p] * 5
    none5 = [None] * 5
    assert L.cvo_device_math_check(-1, None, None, None, 0, *none5) == 1
    assert L.cvo_device_math_check(0, None, None, None, -3, *none5) == 1
    assert L.cvo_device_math_check(3, None, None, None, 0, *none5) == 1
    assert L.cvo_device_math_check(0, None, None, None, 2, *none5) == 1
    assert L.cvo_device_math_check(0, None, None, None, 0, *none5) == 0


def test_abi_struct_sizes():
    """The ctypes mirror has the C layout (checked against field offsets)."""
    from rkhs_ba_amd._abi import CvoParams, CloudView, Assoc, IterRecord
    assert C.sizeof(CvoParams) % 8 == 0
    assert CvoParams.dl.offset % 8 == 0
    assert C.sizeof(CloudView) == 4 * 3 + 4 + 8 * 4  # three int32 + pad + four pointers
    assert C.sizeof(Assoc) == 4 + 4 + 8 + 3 * 8
    assert C.sizeof(IterRecord) == 32
