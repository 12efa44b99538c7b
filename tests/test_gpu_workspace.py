"""The context's device arena (the reference's per-call CvoState,
include/UnifiedCvo/cvo/CvoState.cuh:22-82, allocated on every align) and the
value-ordered association rows of multiframe_is_sorting_inner_product
(src/cvo/CvoGPU.cu:1080-1095).

Bars: repeated calls of one shape do not grow the arena; a workspace cap
below the need fails cleanly with CVO_ERR_LIMIT (9) before allocating, and the
same context then registers bit-identically to an uncapped one; a released
arena regrows to the same size; value-ordered export rows hold the same
(id, value) set as the id-ordered rows, in non-increasing value order (ties
by ascending id), only for rows longer than the threshold.
"""
import os
import sys

import numpy as np
import pytest

from conftest import CONFIGS, params_from_yaml
from rkhs_ba_amd import CvoError, CvoGPU, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_arena_is_reused_capped_and_released():
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, _ = synth.planes_spheres_pair(0)
    c = CvoGPU(params=p)
    try:
        ds, dt = c.upload(src), c.upload(tgt)
        say("align 5k x3 on one context")
        r0 = c.align(ds, dt, np.eye(4))
        used = c.workspace_bytes()
        assert used > 0
        for _ in range(2):
            r = c.align(ds, dt, np.eye(4))
            assert c.workspace_bytes() == used
            np.testing.assert_array_equal(r[1], r0[1])
        # the pair store dominates: N x nearest_neighbors_max x 8 bytes
        assert used >= src.n * p.nearest_neighbors_max * 8
        say("release and regrow")
        c.release_workspace()
        assert c.workspace_bytes() == 0
        r = c.align(ds, dt, np.eye(4))
        assert c.workspace_bytes() == used
        np.testing.assert_array_equal(r[1], r0[1])
    finally:
        c.close()
    say("capped context")
    c2 = CvoGPU(params=p)
    try:
        ds, dt = c2.upload(src), c2.upload(tgt)
        c2.set_workspace_limit(used // 2)
        with pytest.raises(CvoError) as ei:
            c2.align(ds, dt, np.eye(4))
        assert ei.value.status == 9
        assert c2.workspace_bytes() <= used // 2
        c2.set_workspace_limit(used)
        r = c2.align(ds, dt, np.eye(4))
        np.testing.assert_array_equal(r[1], r0[1])
        assert c2.workspace_bytes() <= used
    finally:
        c2.close()


def test_value_ordered_association_rows():
    p = params_from_yaml("planes_spheres_5k.yaml")
    src, tgt, Tgt = synth.planes_spheres_pair(0)
    c = CvoGPU(params=p)
    try:
        say("association by id")
        a = c.compute_association_gpu(src, tgt, Tgt, 0.3)
        thr = 3
        p.multiframe_is_sorting_inner_product = thr
        c.write_params(p)
        say("association by value above 3 pairs")
        b = c.compute_association_gpu(src, tgt, Tgt, 0.3)
    finally:
        c.close()
    np.testing.assert_array_equal(a.row_ptr, b.row_ptr)
    n_sorted = 0
    for i in range(a.rows):
        s, e = a.row_ptr[i], a.row_ptr[i + 1]
        ca, va, cb, vb = a.col[s:e], a.val[s:e], b.col[s:e], b.val[s:e]
        assert sorted(zip(ca.tolist(), va.tolist())) == sorted(zip(cb.tolist(), vb.tolist()))
        if e - s > thr:
            n_sorted += 1
            key = list(zip((-vb).tolist(), cb.tolist()))
            assert key == sorted(key)
        else:
            assert (np.diff(cb) > 0).all()
    assert n_sorted > 100
