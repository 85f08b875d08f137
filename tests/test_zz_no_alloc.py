"""pyr_sfilter_noalloc (fbar with its field temporary in a file-scope device
array) against pyr_sfilter_f64 / pyr_sfilter_r32 through the C-ABI: identical
results on two calls in a row, the second on a smaller grid, no device memory
taken by the calls, and a clear error for a field above the 64 MiB limit."""
import ctypes
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyranda_amd import _lib
    _lib.lib()
    return torch.device("cuda:0")


def _ctx(nn, dev):
    from pyranda_amd.parcop import ParcopContext
    L = [2 * np.pi * (n - 1) / n for n in nn]
    return ParcopContext(nn, (0.0, 0.0, 0.0), L, ["PERI", "PERI", "NONE", "NONE", "PERI", "PERI"], device=dev)


def _field(nn, dev, dtype, seed):
    g = torch.Generator().manual_seed(seed)
    t = torch.rand(tuple(reversed(nn)), generator=g, dtype=torch.float64) + 0.5
    return t.to(dev, dtype).permute(2, 1, 0)


def _noalloc(ctx, v, out):
    from pyranda_amd import _lib
    ctx._stream()
    prec = _lib.F64 if v.dtype == torch.float64 else _lib.R32
    return ctx._L.pyr_sfilter_noalloc(ctx._h, prec, ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(out.data_ptr()))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sfilter_noalloc_equals_sfilter(dev, dtype):
    from pyranda_amd import _lib
    big, small = (96, 80, 64), (40, 33, 20)
    _say("noalloc fbar", dtype, big, "then", small)
    ctxs = {nn: _ctx(nn, dev) for nn in (big, small)}
    ins = {nn: _field(nn, dev, dtype, i + 3) for i, nn in enumerate((big, small))}
    want = {nn: ctxs[nn].sfilter(ins[nn]) for nn in (big, small)}   # the allocating entry point
    outs = {nn: torch.empty_like(want[nn]) for nn in (big, small)}
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(dev)[0]
    for nn in (big, small):   # two calls in a row, the second smaller
        rc = _noalloc(ctxs[nn], ins[nn], outs[nn])
        assert rc == 0, _lib.lib().pyr_last_error().decode()
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(dev)[0]
    # the allocating filter's temporary of the big grid is 3.9 MB (FP64): no
    # device memory was taken by the two calls (measured before any torch
    # comparison could allocate)
    assert free0 - free1 < (1 << 20), (free0, free1)
    for nn in (big, small):
        assert torch.equal(outs[nn], want[nn]), nn
    # the second call's smaller input left the rest of the array as it was;
    # a fresh call on the big grid still gives the same field
    outs[big].zero_()
    assert _noalloc(ctxs[big], ins[big], outs[big]) == 0
    torch.cuda.synchronize()
    assert torch.equal(outs[big], want[big])
    for c in ctxs.values():
        c.close()


def test_sfilter_noalloc_rejects_fields_above_limit(dev):
    from pyranda_amd import _lib
    nn = (256, 256, 130)   # 8.5M FP64 points: 68 MB > 64 MiB
    ctx = _ctx(nn, dev)
    v = torch.zeros(1, dtype=torch.float64, device=dev)   # never read: the size check comes first
    rc = ctx._L.pyr_sfilter_noalloc(ctx._h, _lib.F64, ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(v.data_ptr()))
    assert rc != 0
    msg = _lib.lib().pyr_last_error().decode()
    _say("rejected:", msg)
    assert "64 MiB" in msg
    ctx.close()
