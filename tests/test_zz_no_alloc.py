"""dvm_bs_self_noalloc (the allocation-free blob-on-blob sum) against
dvm_bs_self through the C-ABI: two calls in a row, the second smaller (the
scratch array keeps the first call's contents), in both arithmetic forms, and
a size beyond the scratch array rejected before anything runs."""
import ctypes

import numpy as np
import pytest

import dvmpp_amd as dvm
from conftest import config2_blobs, rel_err, to_dev
from dvmpp_amd._lib import DvmError, check, load

pytestmark = pytest.mark.gpu


def _noalloc(ctx, prec, X, Z, C, S):
    import torch
    u = torch.empty_like(X)
    w = torch.empty_like(X)
    ptr = [ctypes.c_void_p(t.data_ptr()) for t in (X, Z, C, S, u, w)]
    check(ctx.h, load().dvm_bs_self_noalloc(ctx.h, prec, X.numel(), *ptr))
    return u, w


def test_noalloc_matches_bs_self(gpu_ctx):
    import torch
    assert load().dvm_bs_self_noalloc_scratch_bytes() == 64 << 20
    for n in (100_000, 20_000):  # second call smaller
        X, Z, C, S = to_dev(*config2_blobs(n, seed=n))
        u64, w64 = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F64)
        u32, w32 = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F32C)
        nu64, nw64 = _noalloc(gpu_ctx, dvm.F64, X, Z, C, S)
        nu32, nw32 = _noalloc(gpu_ctx, dvm.F32C, X, Z, C, S)
        torch.cuda.synchronize()
        a64 = np.concatenate([u64.cpu().numpy(), w64.cpu().numpy()])
        b64 = np.concatenate([nu64.cpu().numpy(), nw64.cpu().numpy()])
        a32 = np.concatenate([u32.cpu().numpy(), w32.cpu().numpy()])
        b32 = np.concatenate([nu32.cpu().numpy(), nw32.cpu().numpy()])
        assert rel_err(b64, a64) <= 1e-12, n      # FP64 bar: 1e-10
        assert rel_err(b32, a64) <= 1e-5, n       # FP32C bar against FP64
        assert rel_err(b32, a32) <= 1e-6, n
        if n == 100_000:
            # same chunking as dvm_bs_self at this size: bit for bit
            assert np.array_equal(b64, a64)


def test_noalloc_rejects_beyond_scratch(gpu_ctx):
    import torch
    n = 2_000_000
    X, Z, C, S = to_dev(*config2_blobs(n, seed=3))
    with pytest.raises(DvmError) as e:
        _noalloc(gpu_ctx, dvm.F32C, X, Z, C, S)
    assert e.value.status == 2 and "scratch" in str(e.value)
    with pytest.raises(DvmError) as e:
        _noalloc(gpu_ctx, dvm.FMM, X, Z, C, S)
    assert e.value.status == 1
    torch.cuda.synchronize()
    # the context still works, and so does the entry point at a size that fits
    m = 4096
    Xs, Zs, Cs, Ss = (t[:m].contiguous() for t in (X, Z, C, S))
    ref = dvm.bs_self(gpu_ctx, Xs, Zs, Cs, Ss, prec=dvm.F64)
    got = _noalloc(gpu_ctx, dvm.F64, Xs, Zs, Cs, Ss)
    torch.cuda.synchronize()
    assert rel_err(got[0].cpu().numpy(), ref[0].cpu().numpy()) <= 1e-12
