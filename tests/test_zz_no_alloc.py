"""pm_disparity_planes_noalloc (include/pm_io.h) against
pm_disparity_device, through the C ABI.

The entry point runs pm_disparity_device's kernels on caller-owned plane
fields with the LR-checked field in the library's fixed scratch array, so
its disparity must equal the context's bit for bit for every
post-processing mode; two calls in a row, the second on a smaller pair,
check that the shared array carries nothing from one call to the next.
"""
import ctypes

import numpy as np
import pytest

import cvt_amd
from cvt_amd import _native
from test_gpu_parity import first_diff, same

pytestmark = pytest.mark.gpu

MAX_PIXELS = 1 << 21
PPS = [cvt_amd.DISPARITY_PP_NONE, cvt_amd.DISPARITY_PP_LRCHECK,
       cvt_amd.DISPARITY_PP_LRCHECK_FILL]


def noalloc(dL, dR, W, H, pp, out, maxdispdiff=1.0, maxdeg=5.0,
            stream=None):
    L = _native.lib()
    return L.pm_disparity_planes_noalloc(
        ctypes.c_void_p(dL), ctypes.c_void_p(dR), W, H, pp, maxdispdiff,
        maxdeg, ctypes.c_void_p(out), 0, stream)


def context_run(synth, W, H, seed):
    """plane fields (device tensors) and the context's disparity per pp"""
    import torch
    L, R, _ = synth.make_pair(W, H, cells=5, dmax=16.0, channels=4,
                              seed=seed)
    want = {}
    with cvt_amd.Context(radius=3, disparity_max=16.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(2)
        planes = [torch.from_numpy(ctx.planes(v)).cuda() for v in (0, 1)]
        for pp in PPS:
            d = torch.empty((H, W), device="cuda")
            ctx.disparity_device(d.data_ptr(), pp, 1.0, 5.0)
            ctx.sync()
            want[pp] = d.cpu().numpy()
    return planes, want


def test_noalloc_equals_disparity_device_twice(synth):
    import torch
    print("no-alloc disparity: 157 x 83 then 61 x 37", flush=True)
    runs = [(157, 83, 21), (61, 37, 22)]      # the second one smaller
    cases = [context_run(synth, W, H, seed) for W, H, seed in runs]
    stream = torch.cuda.Stream()
    for (W, H, _), (planes, want) in zip(runs, cases):
        for pp in PPS:
            out = torch.full((H, W), -7.0, device="cuda")
            torch.cuda.synchronize()
            rc = noalloc(planes[0].data_ptr(), planes[1].data_ptr(), W, H, pp,
                         out.data_ptr(), stream=stream.cuda_stream)
            assert rc == _native.PM_OK, _native.lib().pm_io_last_error()
            stream.synchronize()
            got = out.cpu().numpy()
            assert same(got, want[pp]), (W, H, pp,
                                         first_diff(got, want[pp]))


def test_noalloc_refuses_sizes_beyond_the_scratch():
    import torch
    buf = torch.zeros(16, device="cuda")
    p = buf.data_ptr()
    rc = noalloc(p, p, 2048, 1025, cvt_amd.DISPARITY_PP_LRCHECK, p)
    assert rc == _native.PM_ERR_ARG
    assert b"static scratch" in _native.lib().pm_io_last_error()
    assert noalloc(p, p, 0, 4, 1, p) == _native.PM_ERR_ARG
    assert noalloc(p, p, 4, 4, 7, p) == _native.PM_ERR_ARG
    assert noalloc(None, p, 4, 4, 1, p) == _native.PM_ERR_ARG
