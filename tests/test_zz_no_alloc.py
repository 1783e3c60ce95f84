"""sptc_index_match_noalloc: the index match with its working memory in a
fixed device array (no allocation), against sptc_index_match on the same
device inputs -- two calls in a row, the second smaller, for the direct
and the distinct-key contract table -- and its size limit."""
import ctypes

import numpy as np
import pytest

import oracle as orc
from gsparc_amd import sptc
from helpers import random_coo

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(a.view(np.int64)).cuda()


def _host(t):
    return t.cpu().numpy().view(np.uint64)


def _keys(rng, dims, nnz, cx, cy):
    X = random_coo(rng, dims, nnz)
    Y = random_coo(rng, dims, int(nnz * 0.8))
    cw = orc.common_cwidth(X.dims, Y.dims, cx, cy)
    xl = orc.layout(X.dims, cx, 1, cw)
    yl = orc.layout(Y.dims, cy, 2, cw)
    xk, _ = orc.sort_pairs(orc.linearize(X.idx, xl[0], xl[1]), X.val)
    yk, _ = orc.sort_pairs(orc.linearize(Y.idx, yl[0], yl[1]), Y.val)
    return xk, yk, xl[4], yl[3]


def _match(torch, fn, xk, yk, ncb, ynf):
    dx, dy = _dev(torch, xk), _dev(torch, yk)
    nx = len(xk)
    out = [torch.full((n,), -1, dtype=torch.int64, device="cuda") for n in (nx, nx, nx + 1)]
    ir = ctypes.c_uint64()
    st = fn(ctypes.c_void_p(dx.data_ptr()), nx, ctypes.c_void_p(dy.data_ptr()), len(yk), ncb,
            ynf, *[ctypes.c_void_p(o.data_ptr()) for o in out], ctypes.byref(ir),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    sptc.check(st)
    torch.cuda.synchronize()
    return [_host(o) for o in out], ir.value


@pytest.mark.parametrize("dims,cx,cy", [((200, 150, 120), [2], [0]),          # direct table
                                        ((300, 2000, 4000), [1, 2], [1, 2])])  # 23 bits: search
def test_noalloc_equals_index_match(gpu, dims, cx, cy):
    import torch
    L = sptc.lib()
    L.sptc_index_match_noalloc.argtypes = L.sptc_index_match.argtypes
    rng = np.random.default_rng(sum(dims) + len(cx))
    for nnz in (300_000, 40_000):  # second call smaller: stale workspace must not leak in
        xk, yk, ncb, ynf = _keys(rng, dims, nnz, cx, cy)
        want, ir_want = _match(torch, L.sptc_index_match, xk, yk, ncb, ynf)
        before = sptc.pool_stats()
        got, ir_got = _match(torch, L.sptc_index_match_noalloc, xk, yk, ncb, ynf)
        after = sptc.pool_stats()
        assert after["hip_mallocs"] == before["hip_mallocs"]
        assert after["live_bytes"] == before["live_bytes"]
        assert ir_got == ir_want > 0
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)


def test_noalloc_rejects_sizes_above_its_workspace(gpu):
    L = sptc.lib()
    L.sptc_index_match_noalloc.argtypes = L.sptc_index_match.argtypes
    ir = ctypes.c_uint64()
    # 30-bit contract field with 2^21 Y nonzeros: the distinct-key table
    # would need 32 MiB + 16 B; rejected before any device work
    st = L.sptc_index_match_noalloc(ctypes.c_void_p(8), 1, ctypes.c_void_p(8), 1 << 21, 30, 10,
                                    ctypes.c_void_p(8), ctypes.c_void_p(8),
                                    ctypes.c_void_p(8), ctypes.byref(ir), None)
    assert st == sptc.ERR_BAD_ARGS
    assert b"fixed workspace" in L.sptc_last_error()
