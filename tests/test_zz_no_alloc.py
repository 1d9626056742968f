"""zj_gpu_residual_norm_noalloc (the allocation-free variant of
zj_gpu_residual_norm) against the allocating entry point, through the C ABI,
on two calls in a row (the second with a smaller input), and its size limit.
Runs on an MI355X."""
import math

import numpy as np
import pytest

from zhijian_amd import gpu_ops as G
from zhijian_amd._lib import ZJError
from zhijian_amd.hiprt import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def test_residual_norm_noalloc_matches_allocating():
    rng = np.random.default_rng(2)
    stream = Stream()
    for n_elem, n_sp in ((65536, 25), (700, 16), (1, 4)):  # then a smaller input, then the smallest
        print(f"case: residual norm, n_elem {n_elem}, n_sp {n_sp}", flush=True)
        R = rng.standard_normal(n_elem * n_sp * 4) * 10.0 ** rng.uniform(-3, 3)
        d = DeviceArray(R)
        a = G.compute_residual_norm(d.ptr, n_elem, n_sp)
        b = G.compute_residual_norm_noalloc(d.ptr, n_elem, n_sp)
        c = G.compute_residual_norm_noalloc(d.ptr, n_elem, n_sp, stream.ptr)
        assert a == b == c
        assert abs(b - math.sqrt(math.fsum(R * R))) <= 1e-14 * b
    assert G.compute_residual_norm_noalloc(None, 0, 4) == 0.0


def test_residual_norm_noalloc_size_limit():
    with pytest.raises(ZJError) as e:
        G.compute_residual_norm_noalloc(None, 16777217, 1)  # 67108868 values
    assert e.value.code == 1 and "maximum" in str(e.value)
