"""i360hip_compute_lab_stats_dev_noalloc (the statistics entry point that
allocates nothing, its scratch a static device array) against
i360hip_compute_lab_stats_dev through the C ABI: bit-identical means and
deviations on two calls in a row, the second on a smaller image (the array
keeps the first call's contents, which the second must fully replace)."""
import ctypes as C

import numpy as np
import pytest

from image360_amd import _abi
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


def _stats(L, fn, ctx, img):
    h, w = img.shape[:2]
    m = (C.c_double * 3)()
    s = (C.c_double * 3)()
    rc = fn(ctx, C.c_void_p(img.data_ptr()), w, h, m, s, None)
    assert rc == 0, (rc, (L.i360hip_last_error(ctx) or b"").decode())
    return np.array(list(m) + list(s))


def test_noalloc_statistics_equal_the_allocating_entry_point():
    import torch

    L = _abi.lib()
    ctx = C.c_void_p()
    assert L.i360hip_create(0, C.byref(ctx)) == 0, (L.i360hip_last_error(None) or b"").decode()
    try:
        g = torch.Generator(device="cuda").manual_seed(41)
        for (h, w) in ((1531, 2077), (97, 133)):
            note(f"no-alloc lab statistics {w}x{h}")
            img = torch.randint(0, 65536, (h, w, 3), generator=g, device="cuda", dtype=torch.int32)
            img = img.to(torch.int16).contiguous()
            torch.cuda.synchronize()
            # the allocation-free call first: on the first image it runs
            # before the context has allocated any colour scratch
            got = _stats(L, L.i360hip_compute_lab_stats_dev_noalloc, ctx, img)
            want = _stats(L, L.i360hip_compute_lab_stats_dev, ctx, img)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
            assert 0.0 < want[0] < 100.0 and (want[3:] > 0).all()
        # argument checks as the allocating entry point
        m = (C.c_double * 3)()
        s = (C.c_double * 3)()
        assert L.i360hip_compute_lab_stats_dev_noalloc(ctx, None, 4, 4, m, s, None) == _abi.INVALID_PARAM
        assert L.i360hip_compute_lab_stats_dev_noalloc(ctx, C.c_void_p(img.data_ptr()), 0, 4, m, s, None) == \
            _abi.INVALID_PARAM
    finally:
        L.i360hip_destroy(ctx)
