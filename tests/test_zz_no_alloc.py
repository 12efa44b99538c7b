"""The no-allocation inner product (cvo_inner_product_d_noalloc) through the C
ABI against the allocating one (cvo_inner_product_d): two calls in a row, the
second on smaller clouds, each within 1e-12 relative of the allocating entry
point on the same input (same kernels; the scratch grid is capped at 2^21
cells, which may change a row's FP64 summation order only); a cloud beyond the
documented maximum is rejected with CVO_ERR_LIMIT; the context's arena is
neither grown nor disturbed by the call."""
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


def test_noalloc_inner_product_matches_the_allocating_one():
    c = CvoGPU(os.path.join(CONFIGS, "street_60k.yaml"))
    try:
        p = params_from_yaml("street_60k.yaml")
        c.write_params(p)
        cases = [(synth.street_pair(1, n=20000), 0.3), (synth.street_pair(2, n=3000), 0.5)]
        for (src, tgt, Tgt), ell in cases:
            a, b = c.upload(src), c.upload(tgt)
            say(f"noalloc inner product n={src.n} ell={ell}")
            ref = c.inner_product_gpu(a, b, Tgt, ell)
            ws = c.workspace_bytes()
            got = c.inner_product_noalloc(a, b, Tgt, ell)
            assert c.workspace_bytes() == ws
            assert ref > 0 and abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
            # the arena is untouched: the allocating call gives the same again
            assert c.inner_product_gpu(a, b, Tgt, ell) == ref
            a.free()
            b.free()
        say("noalloc inner product beyond the maximum")
        big = synth.Cloud(np.random.default_rng(0).uniform(-5, 5, size=(131073, 3)).astype(np.float32),
                          np.zeros((131073, 5), np.float32), np.full((131073, 19), 1.0 / 19, np.float32))
        da = c.upload(big)
        with pytest.raises(CvoError, match="size beyond a documented limit"):
            c.inner_product_noalloc(da, da, np.eye(4), 0.5)
        da.free()
    finally:
        c.close()
