"""GPU tests of the planned form (sptc_plan / sptc_execute /
sptc_partition_plan): operands prepared once, executed repeatedly, with the
same results as the one-call path, and the dynamic chunk plan obeying the
reference's Dynamic_partition invariants (contraction.cuh:143-245): chunks
cover the X rows in order, never split a row, and sum to the product count."""
import numpy as np
import pytest

from gsparc_amd import sptc
from helpers import REL_TOL, assert_same_coo, load_fixtures, random_coo

pytestmark = pytest.mark.gpu
FIXTURES = load_fixtures()


def _same(a, b):
    for x, y in zip(a.to_host(), b.to_host()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_plan_execute_matches_golden(gpu, case):
    p = sptc.Plan(case["X"], case["Y"], case["cx"], case["cy"])
    for _ in range(2):
        r = p.execute()
        idx, v = r.decode()
        assert_same_coo([a.astype(np.int64) for a in idx], v, case["zcoords"], case["zval"],
                        rtol=REL_TOL)


@pytest.mark.parametrize("dense", [0, -1])
@pytest.mark.parametrize("seed", [3, 4])
def test_plan_execute_equals_contract_bitwise(gpu, seed, dense):
    rng = np.random.default_rng(seed)
    X = random_coo(rng, (70, 40, 30), 30_000, positive=bool(seed % 2))
    Y = random_coo(rng, (30, 90, 25), 25_000, positive=bool(seed % 2))
    one = sptc.SpTC(X, Y, [2], [0], dense_accumulator=dense)
    p = sptc.Plan(X, Y, [2], [0], dense_accumulator=dense)
    info = p.info
    assert info["x_nnz"] == X.nnz and info["y_nnz"] == Y.nnz
    assert info["out_bits"] == one.info.x_nfbits + one.info.y_nfbits
    assert info["device_bytes"] >= 12 * (X.nnz + Y.nnz)
    for _ in range(3):
        r = p.execute()
        assert r.nnz == one.nnz and r.ir_nnz == one.ir_nnz
        assert r.info.accumulator == one.info.accumulator
        _same(r, one)
        np.testing.assert_array_equal(r.row_keys(), one.row_keys())
        assert r.info.phase_ms[1] < 1.0  # operands were prepared by the plan


def test_partition_plan_invariants_and_budget_invariance(gpu):
    rng = np.random.default_rng(12)
    X = random_coo(rng, (60, 50, 40), 60_000)
    full = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1)
    p = sptc.Plan(X, X, [2], [2], dense_accumulator=-1)
    rows, prods, ir = p.partitions()
    assert ir == full.ir_nnz
    assert len(rows) == 2 and rows[0] == 0 and rows[-1] == full.info.nrows
    budget = full.nnz * 16 + full.ir_nnz * 8 + (128 << 20)
    rows, prods, ir = p.partitions(budget)
    n = len(rows) - 1
    assert n > 1
    assert rows[0] == 0 and rows[-1] == full.info.nrows
    assert prods[0] == 0 and prods[-1] == ir == full.ir_nnz
    assert np.all(np.diff(rows.astype(np.int64)) > 0)
    assert np.all(np.diff(prods.astype(np.int64)) >= 0)
    # chunk boundaries sit on row boundaries: the products before boundary
    # row r are exactly those of rows [0, r) (rows = distinct X free keys in
    # ascending order, products of a nonzero = Y's nonzeros with its
    # contract index)
    cnt_y = np.bincount(X.idx[2].astype(np.int64), minlength=X.dims[2])
    fkey = X.idx[0].astype(np.int64) * X.dims[1] + X.idx[1].astype(np.int64)
    _, inv = np.unique(fkey, return_inverse=True)
    per_row = np.bincount(inv, weights=cnt_y[X.idx[2].astype(np.int64)]).astype(np.int64)
    cum = np.concatenate([[0], np.cumsum(per_row)])
    np.testing.assert_array_equal(cum[rows.astype(np.int64)], prods.astype(np.int64))
    # executing under a budget uses the chunk plan sptc_partition_plan
    # reports for it (same device memory in use), and the output does not
    # depend on the chunking
    pb = sptc.Plan(X, X, [2], [2], dense_accumulator=-1, mem_budget_bytes=budget)
    rows_b, prods_b, _ = pb.partitions()
    r = pb.execute()
    assert r.info.nchunks == len(rows_b) - 1 > 1
    _same(r, full)


def test_partition_plan_rejects_a_budget_below_one_row(gpu):
    rng = np.random.default_rng(13)
    X = random_coo(rng, (20, 30, 40), 20_000)
    p = sptc.Plan(X, X, [2], [2], dense_accumulator=-1)
    with pytest.raises(sptc.SpTCError) as e:
        p.partitions(1 << 20)
    assert e.value.status == sptc.ERR_OOM_BUDGET


def test_plan_errors(gpu):
    rng = np.random.default_rng(14)
    X = random_coo(rng, (10, 10, 10), 50)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.Plan(X, X, [5], [2])
    assert e.value.status == sptc.ERR_BAD_ARGS


def test_execute_options_per_call(gpu):
    """As in the reference (the accumulator is an argument of SpTC, not of
    the operand preparation), one plan executes with either accumulator."""
    rng = np.random.default_rng(15)
    X = random_coo(rng, (40, 12, 30), 9_000)
    p = sptc.Plan(X, X, [2], [2])
    for dense in (1, -1, 0):
        r = p.execute(dense_accumulator=dense)
        one = sptc.SpTC(X, X, [2], [2], dense_accumulator=dense)
        assert r.info.accumulator == one.info.accumulator
        _same(r, one)
    cnt = p.execute(count_only=True)
    assert cnt.nnz == one.nnz and cnt.ir_nnz == one.ir_nnz


def test_accumulate_entry_points(gpu):
    """sptc_accumulate_esc / sptc_accumulate_dense (ESC_Contraction /
    Dense_Contraction on a plan's operands) are the forced accumulators of
    sptc_execute: the LDS row for a narrow column space, the tiled form for
    a wide one."""
    rng = np.random.default_rng(16)
    narrow = random_coo(rng, (40, 12, 30), 9_000)
    wide = random_coo(rng, (300, 200, 40), 30_000)
    for X, dense_kind in ((narrow, 1), (wide, 2)):
        p = sptc.Plan(X, X, [2], [2])
        e, d = p.accumulate_esc(), p.accumulate_dense()
        assert e.info.accumulator == 0 and d.info.accumulator == dense_kind
        _same(e, p.execute(dense_accumulator=-1))
        if dense_kind == 1:
            _same(d, p.execute(dense_accumulator=1))
        else:  # FP64 LDS adds in no fixed order: equal up to the last FP32 bit
            for a, b in zip(d.to_host(), p.execute(dense_accumulator=1).to_host()):
                np.testing.assert_allclose(a.astype(np.float64), b.astype(np.float64),
                                           rtol=2.0 ** -23, atol=0)
        xe, ye, _ = e.to_host()
        xd, yd, _ = d.to_host()
        np.testing.assert_array_equal(xe, xd)
        np.testing.assert_array_equal(ye, yd)
