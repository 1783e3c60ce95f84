"""Static operand partitioning on the GPU.

Operands whose device footprint exceeds half the memory budget are split
into free-key ranges (the reference's FindPartitionNum / Partition /
convert_and_sort_64, tensor_manager.tpp:32-61,633-798) and contracted pair
by pair (the X x Y partition loop, sptc.cuh:42-81).  The result must be the
unpartitioned one entry for entry and bit for bit: a row's products come
from the same X nonzeros in the same order, and each output column from
exactly one Y part.
"""
import numpy as np
import pytest

from gsparc_amd import sptc
from gsparc_amd.coo import COOTensor
from helpers import compare_with_oracle, x_free_keys

pytestmark = pytest.mark.gpu

MB = 1 << 20


def _draw(rng, dims, nnz, hi=None):
    """nnz coordinates drawn uniformly (mode m below hi[m] when given;
    duplicates kept, as the reader keeps them), values in (0.05, 1.05)."""
    hi = hi or dims
    idx = [rng.integers(0, h, nnz, dtype=np.uint32) for h in hi]
    return COOTensor(tuple(dims), idx, (rng.random(nnz) + 0.05).astype(np.float32))


def _operands():
    rng = np.random.default_rng(2024)
    # X: 1e6 nonzeros whose contract index spans 1e6 values; Y: 2e6
    # nonzeros with contract indices below 1e5 -- a tenth of X matches, so
    # the operands are large against the output
    X = _draw(rng, (2000, 3000, 1_000_000), 1_000_000)
    Y = _draw(rng, (1_000_000, 2500, 1500), 2_000_000, hi=(100_000, 2500, 1500))
    return X, Y


def _same(a, b):
    assert a.nnz == b.nnz and a.ir_nnz == b.ir_nnz
    for x, y in zip(a.to_host(), b.to_host()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.row_keys(), b.row_keys())


@pytest.fixture(scope="module")
def operands():
    return _operands()


@pytest.fixture(scope="module")
def whole(operands):
    X, Y = operands
    return sptc.SpTC(X, Y, [2], [0])


@pytest.mark.parametrize("budget_mb,parts", [(200, (2, 2)), (240, (3, 1))])
def test_partitioned_equals_whole(gpu, operands, whole, budget_mb, parts):
    X, Y = operands
    px, py = sptc.find_partition_num(X.nnz, X.order, Y.nnz, Y.order, 4, budget_mb * MB)
    assert (px, py) == parts
    res = sptc.SpTC(X, Y, [2], [0], mem_budget_bytes=budget_mb * MB)
    assert res.info.nchunks == px
    _same(res, whole)
    emitted, bad = res.verify()
    assert bad == 0
    assert abs(res.info.value_sum - whole.info.value_sum) <= 1e-9 * whole.info.value_sum
    print(f"[static] budget {budget_mb} MB: {px} X parts x {py} Y parts, out {res.nnz}",
          flush=True)


def test_partitioned_plan(gpu, operands, whole):
    X, Y = operands
    budget = 200 * MB
    p = sptc.Plan(X, Y, [2], [0], mem_budget_bytes=budget)
    info = p.info
    px, py = sptc.find_partition_num(X.nnz, X.order, Y.nnz, Y.order, 4, budget)
    assert (info["x_parts"], info["y_parts"]) == (px, py)
    assert info["x_nnz"] == X.nnz and info["y_nnz"] == Y.nnz
    # Partition's output: free-key ranges covering every nonzero, counts
    # matching the free keys of the input
    for role, T, cm, keys in ((1, X, [2], x_free_keys(X, Y, [2], [0])),
                              (2, Y, [0], None)):
        nb, kb = p.static_partitions(role)
        assert len(nb) == (px if role == 1 else py) + 1
        assert nb[0] == 0 and nb[-1] == T.nnz
        assert np.all(np.diff(nb.astype(np.int64)) >= 0)
        assert np.all(np.diff(kb.astype(np.float64)) >= 0)
        if keys is None:  # Y free key: modes 1, 2 (last at bit 0)
            shift = sptc.layout(Y.dims, [0], 2)[0]
            keys = np.zeros(T.nnz, np.uint64)
            for m in (1, 2):
                keys |= T.idx[m].astype(np.uint64) << np.uint64(shift[m])
        counts = np.histogram(keys.astype(np.float64), bins=kb.astype(np.float64))[0]
        np.testing.assert_array_equal(counts, np.diff(nb.astype(np.int64)))
    for _ in range(2):
        _same(p.execute(), whole)
    cnt = p.execute(count_only=True)
    assert cnt.nnz == whole.nnz and cnt.ir_nnz == whole.ir_nnz
    assert abs(cnt.info.value_sum - whole.info.value_sum) <= 1e-9 * whole.info.value_sum


def test_partitioned_rows_against_oracle(gpu, operands):
    X, Y = operands
    res = sptc.SpTC(X, Y, [2], [0], mem_budget_bytes=240 * MB)
    rk = res.row_keys()
    pick = rk[np.linspace(0, len(rk) - 1, 40).astype(np.int64)]
    compare_with_oracle(X, Y, [2], [0], res, x_rows=pick)
