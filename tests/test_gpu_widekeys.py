"""Operands whose packed keys need more than 64 bits (the reference's extra::
path, tensor_manager.tpp:300-406, indexMatch.cuh:120-229, sort.cuh:372-526).

Here the contract field becomes the rank of the contracted coordinates among
those of both operands (order-preserving, so the sort and match see the
reference's order), computed on the device one 64-bit group of modes at a
time.  Checked against the oracle's independent FP64 join on the same
contraction with the contracted tuples relabelled to one small mode (equal
tuples stay equal, so products and outputs do not change).
"""
import numpy as np
import pytest

import oracle as orc
from gsparc_amd import sptc
from gsparc_amd.coo import COOTensor
from helpers import REL_TOL

pytestmark = pytest.mark.gpu

W20 = 1 << 20


def _tensor(rng, dims, nnz, cmodes, pool):
    """Uniform indices, except that the contracted coordinates are drawn
    from `pool` distinct tuples (so contractions have matches)."""
    idx = [rng.integers(0, d, nnz, dtype=np.uint64).astype(np.uint32) for d in dims]
    pick = rng.integers(0, pool, nnz)
    for m in cmodes:
        idx[m] = rng.integers(0, dims[m], pool, dtype=np.uint64).astype(np.uint32)[pick]
    return COOTensor(tuple(dims), idx, (rng.random(nnz) + 0.05).astype(np.float32))


def _free_dims(t, cmodes):
    return [t.dims[m] for m in range(t.order) if m not in cmodes]


def _linear(cols, dims):
    lin = np.zeros(len(cols[0]) if cols else 0, np.uint64)
    for c, d in zip(cols, dims):
        lin = lin * np.uint64(d) + c.astype(np.uint64)
    return lin


def _relabel(X, Y, cx, cy):
    """The same contraction with one contracted mode: each distinct tuple
    of contracted coordinates (over both operands) replaced by its id --
    equal tuples stay equal, so the products and outputs are unchanged.
    Lets the FP64 join's mixed-radix contract ids stay below 2^64."""
    tx = np.stack([X.idx[m].astype(np.uint64) for m in cx], 1)
    ty = np.stack([Y.idx[m].astype(np.uint64) for m in cy], 1)
    _, inv = np.unique(np.concatenate([tx, ty]), axis=0, return_inverse=True)
    inv = inv.reshape(-1).astype(np.uint32)
    K = int(inv.max()) + 1

    def one(t, cm, ids):
        free = [m for m in range(t.order) if m not in cm]
        return COOTensor(tuple([t.dims[m] for m in free] + [K]),
                         [t.idx[m] for m in free] + [ids], t.val)
    return one(X, cx, inv[:X.nnz]), one(Y, cy, inv[X.nnz:])


def _check(res, X, Y, cx, cy):
    Xr, Yr = _relabel(X, Y, cx, cy)
    xf, yf, v = orc.join_fp64(Xr, Yr, [Xr.order - 1], [Yr.order - 1])
    idx, vals = res.decode()
    fx, fy = _free_dims(X, cx), _free_dims(Y, cy)
    nx = len(fx)
    gx = _linear(idx[:nx], fx)
    gy = _linear(idx[nx:], fy)
    assert res.nnz == len(v) > 0
    np.testing.assert_array_equal(gx, xf)
    np.testing.assert_array_equal(gy, yf)
    rel = np.abs(vals.astype(np.float64) - v) / np.abs(v)
    assert rel.max() <= REL_TOL
    return rel.max()


def test_order5_wide_self_contraction(gpu):
    """78-bit keys: order 5, dims 3000 x 2^20 x 2^20 x 50 x 2^20, modes
    {1, 2} (40 bits) contracted."""
    rng = np.random.default_rng(71)
    dims = (3000, W20, W20, 50, W20)
    X = _tensor(rng, dims, 100_000, (1, 2), 3000)
    with pytest.raises(sptc.SpTCError):
        sptc.layout(dims, [1, 2], 1)  # the plain layout does not fit 64 bits
    res = sptc.SpTC(X, X, [1, 2], [1, 2])
    assert res.info.ncbits == int(np.ceil(np.log2(3000)))  # rank of <= 3000 tuples
    _check(res, X, X, [1, 2], [1, 2])
    # the dense accumulator cannot apply (column space); ESC and the planned
    # form give the same result bit for bit
    plan = sptc.Plan(X, X, [1, 2], [1, 2])
    for a, b in zip(plan.execute().to_host(), res.to_host()):
        np.testing.assert_array_equal(a, b)


def test_wide_contract_field_above_64_bits(gpu):
    """Four contracted modes of 20 bits (80 bits of contract field): ranked
    in two groups (60 bits, then the rank above the last 20)."""
    rng = np.random.default_rng(72)
    dims = (100, W20, W20, W20, W20, 60)
    X = _tensor(rng, dims, 60_000, (1, 2, 3, 4), 2000)
    Y = _tensor(rng, (W20, W20, W20, W20, 40), 50_000, (0, 1, 2, 3), 2000)
    # share the tuples: Y's contracted coordinates drawn from X's
    sel = rng.integers(0, X.nnz, Y.nnz)
    Y = COOTensor(Y.dims, [X.idx[1][sel], X.idx[2][sel], X.idx[3][sel], X.idx[4][sel],
                           Y.idx[4]], Y.val)
    res = sptc.SpTC(X, Y, [1, 2, 3, 4], [0, 1, 2, 3])
    _check(res, X, Y, [1, 2, 3, 4], [0, 1, 2, 3])


def test_wide_keys_statically_partitioned(gpu):
    """Ranks are taken per partition pair; the result does not change."""
    rng = np.random.default_rng(73)
    dims = (3000, W20, W20, 50, W20)
    X = _tensor(rng, dims, 400_000, (1, 2), 200_000)
    whole = sptc.SpTC(X, X, [1, 2], [1, 2])
    budget = 80 << 20
    px, py = sptc.find_partition_num(X.nnz, X.order, X.nnz, X.order, 4, budget)
    assert px * py > 1
    part = sptc.SpTC(X, X, [1, 2], [1, 2], mem_budget_bytes=budget)
    assert part.nnz == whole.nnz and part.ir_nnz == whole.ir_nnz
    for a, b in zip(part.to_host(), whole.to_host()):
        np.testing.assert_array_equal(a, b)


def test_free_field_above_64_bits_is_reported(gpu):
    rng = np.random.default_rng(74)
    dims = (W20 * 1024, W20 * 1024, W20 * 1024, 10)  # 90 free bits
    X = _tensor(rng, dims, 1000, (3,), 5)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(X, X, [3], [3])
    assert e.value.status == sptc.ERR_KEY_OVERFLOW
