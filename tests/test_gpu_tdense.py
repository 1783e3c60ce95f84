"""The tiled dense accumulator (dense_accumulator 2, or 1 when the column
space exceeds one LDS row): per-row dense accumulation in LDS over column
tiles with FP64 slots, against the CPU oracle -- indices bit-exact, values
within REL_TOL of FP64 (relative to the sum of |products| when signed) --
and against the ESC accumulator (same entries)."""
import numpy as np
import pytest

from gsparc_amd import sptc
from gsparc_amd.coo import COOTensor, UNIFORM, ZIPF, VAL_UNIFORM01, generate
from helpers import compare_with_oracle, random_coo

pytestmark = pytest.mark.gpu


def _same_entries(a, b):
    xa, ya, _ = a.to_host()
    xb, yb, _ = b.to_host()
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(ya, yb)


@pytest.mark.parametrize("seed,positive", [(1, True), (2, False), (3, True)])
def test_random_wide_columns(gpu, seed, positive):
    rng = np.random.default_rng(seed)
    X = random_coo(rng, (90, 70, 30), 25_000, positive=positive)
    Y = random_coo(rng, (30, 600, 500), 40_000, positive=positive)
    res = sptc.SpTC(X, Y, [2], [0], dense_accumulator=2)
    assert res.info.accumulator == 2
    compare_with_oracle(X, Y, [2], [0], res)
    _same_entries(res, sptc.SpTC(X, Y, [2], [0], dense_accumulator=-1))


def test_zipf_rows_many_owners_and_tiles(gpu):
    """Zipf rows: heavy rows with hundreds of owners (several owner batches
    per tile) over a column space of many tiles, among many light rows."""
    X = generate(0x5EED00AB, (400, 300, 2000), (ZIPF, ZIPF, ZIPF), (1.2, 1.2, 1.1), 120_000,
                 VAL_UNIFORM01)
    res = sptc.SpTC(X, X, [2], [2], dense_accumulator=2)
    owners = np.bincount(np.unique(X.idx[0].astype(np.int64) * 300 + X.idx[1],
                                   return_inverse=True)[1])
    assert owners.max() > 512
    rk = res.row_keys()
    pick = np.unique(np.concatenate([rk[:20], rk[-5:], rk[:: max(1, len(rk) // 40)]]))
    compare_with_oracle(X, X, [2], [2], res, x_rows=pick)
    esc = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1)
    _same_entries(res, esc)
    assert abs(res.info.value_sum - esc.info.value_sum) <= 1e-9 * abs(esc.info.value_sum)


def test_duplicate_y_coordinates(gpu):
    rng = np.random.default_rng(4)
    X = random_coo(rng, (50, 40, 20), 8000)
    Y = random_coo(rng, (20, 300, 300), 30_000)
    Y = COOTensor(Y.dims, [np.concatenate([a, a[:3000]]) for a in Y.idx],
                  np.concatenate([Y.val, Y.val[:3000] * 0.25]))
    res = sptc.SpTC(X, Y, [2], [0], dense_accumulator=2)
    compare_with_oracle(X, Y, [2], [0], res)


def test_output_budget(gpu):
    """The whole output stays resident: a budget below it is reported, one
    above it gives the unconstrained result."""
    rng = np.random.default_rng(5)
    X = random_coo(rng, (200, 150, 40), 60_000)
    full = sptc.SpTC(X, X, [2], [2], dense_accumulator=2)
    assert full.info.nchunks == 1
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(X, X, [2], [2], dense_accumulator=2, mem_budget_bytes=full.nnz * 8)
    assert e.value.status == sptc.ERR_OOM_BUDGET
    ok = sptc.SpTC(X, X, [2], [2], dense_accumulator=2,
                   mem_budget_bytes=full.nnz * 16 + (256 << 20))
    for a, b in zip(full.to_host(), ok.to_host()):
        np.testing.assert_array_equal(a, b)
    emitted, bad = ok.verify()
    assert bad == 0
