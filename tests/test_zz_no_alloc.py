"""sortgpu_trace_closest_noalloc (the entry point for callers that cannot allocate) against
sortgpu_trace_closest: every hit bitwise equal, on two sorted batches in a row (the second smaller,
so it runs over what the first left in the shared device array), on a batch below the sorting
threshold, and with the sort forced by the tuning knob; larger batches than the array holds are
refused before any launch; the calls leave the device's free memory unchanged.
"""
import ctypes
import gc

import numpy as np
import pytest
import torch

from sort_amd import Obvh, SortGpuError, lib, scenes
from sort_amd._lib import check
from sort_amd.batch import set_tuning

pytestmark = pytest.mark.gpu

LIMIT = 1 << 22


def _noalloc(acc, d_rays, n=None, out=None):
    n = d_rays.shape[0] if n is None else n
    if out is None:
        out = torch.empty((d_rays.shape[0], 4), dtype=torch.float32, device=d_rays.device)
    s = torch.cuda.current_stream(d_rays.device).cuda_stream
    check(lib().sortgpu_trace_closest_noalloc(acc.handle, d_rays.data_ptr(), n, out.data_ptr(), s), acc.ctx.handle)
    return out


@pytest.fixture(scope="module")
def soup_acc(gpu_ctx):
    pos, idx = scenes.soup(1 << 18, seed=0x5EED0A11)
    acc = Obvh(gpu_ctx)
    acc.build(pos, idx)
    return acc


def test_noalloc_equals_closest_on_successive_batches(gpu_ctx, soup_acc):
    acc = soup_acc
    rays = scenes.random_rays(LIMIT, 0xA110C, lo=-0.2, hi=1.2, tmin=1e-4)
    d_all = torch.from_numpy(rays).to("cuda:0")
    for n, label in [(LIMIT, "full"), (300_000, "smaller"), (1000, "below the sort threshold")]:
        print(f"[noalloc] {label}: {n} rays", flush=True)
        d = d_all[:n].contiguous()
        ref = acc.get_intersect(d).cpu().numpy()
        out = torch.empty((n, 4), dtype=torch.float32, device="cuda:0")  # the caller's output
        gc.collect()  # objects of earlier tests free their device memory now, not during the call
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(0)[0]
        _noalloc(acc, d, out=out)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info(0)[0]
        got = out.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), label
        assert (ref[:, 1].view(np.uint32) != 0xFFFFFFFF).mean() > 0.3
        if n < LIMIT:  # the first call may load the code object that holds the array
            assert free1 >= free0, f"{label}: the call took {free0 - free1} bytes of device memory"
    acc.check_errors()


def test_noalloc_forced_order_and_limit(gpu_ctx, soup_acc):
    acc = soup_acc
    rays = scenes.random_rays(5000, 0xA110D, lo=-0.2, hi=1.2)
    d = torch.from_numpy(rays).to("cuda:0")
    ref = acc.get_intersect(d).cpu().numpy()
    set_tuning(gpu_ctx, reorder=2)  # sort even this small batch
    try:
        print("[noalloc] forced order: 5000 rays", flush=True)
        got = _noalloc(acc, d).cpu().numpy()
    finally:
        set_tuning(gpu_ctx)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # refused on the host before anything is launched or read
    with pytest.raises(SortGpuError) as e:
        _noalloc(acc, d, n=LIMIT + 1)
    assert e.value.code == -1 and "no-allocation limit" in str(e.value)
    assert lib().sortgpu_trace_closest_noalloc(None, None, 1, None, None) == -1
