"""GPU parity: the HIP path (libsptc_hip.so, through its C ABI) against the CPU
oracle and the golden fixtures.  Index work must be bit-exact; FP32 values
within REL_TOL (1e-5) relative error of the FP64 oracle."""
import ctypes

import numpy as np
import pytest

import oracle as orc
from gsparc_amd import sptc
from gsparc_amd.coo import COOTensor
from gsparc_amd.workloads import PLAIN_SMALL
from helpers import (REL_TOL, assert_same_coo, compare_with_oracle, decode_x, decode_y,
                     dense_friendly, load_fixtures, max_rel_err, random_coo)

pytestmark = pytest.mark.gpu
FIXTURES = load_fixtures()


def _dev(torch, a):
    if a.dtype == np.uint64:
        return torch.from_numpy(a.view(np.int64)).cuda()
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def _host_u64(t):
    return t.cpu().numpy().view(np.uint64)


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# ---- (a) linearize ---------------------------------------------------------

@pytest.mark.parametrize("seed", range(4))
def test_linearize_matches_oracle(gpu, seed):
    import torch
    rng = np.random.default_rng(seed)
    order = int(rng.integers(2, 6))
    dims = [int(x) for x in rng.integers(2, 5000, size=order)]
    nc = int(rng.integers(1, order))
    cm = [int(x) for x in rng.permutation(order)[:nc]]
    t = random_coo(rng, dims, 100_000)
    for role in (1, 2):
        sh, wd, fp, nf, ncb = orc.layout(dims, cm, role)
        sh2, wd2, nf2, ncb2 = sptc.layout(dims, cm, role)
        assert (sh2, wd2, nf2, ncb2) == (sh, wd, nf, ncb)
        idx = [_dev(torch, a) for a in t.idx]
        keys = torch.empty(t.nnz, dtype=torch.int64, device="cuda")
        ptrs = (ctypes.c_void_p * order)(*[a.data_ptr() for a in idx])
        sptc.check(sptc.lib().sptc_linearize(order, t.nnz, ptrs, 4, (ctypes.c_int32 * order)(*sh),
                                             (ctypes.c_int32 * order)(*wd),
                                             ctypes.c_void_p(keys.data_ptr()), _stream(torch)))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_host_u64(keys), orc.linearize(t.idx, sh, wd))


# ---- (b) radix sort --------------------------------------------------------

@pytest.mark.parametrize("n,bits,rb", [(0, 8, 8), (1, 3, 8), (4095, 20, 8), (4096, 20, 8),
                                       (4097, 33, 8), (100_000, 43, 8), (1_000_003, 48, 8),
                                       (2_000_000, 64, 8), (300_000, 31, 11),
                                       (1_500_000, 44, 11), (700_000, 36, 9),
                                       (900_001, 41, 10), (3_000_000, 36, 0)])
def test_sort_pairs_stable_bit_exact(gpu, n, bits, rb):
    import torch
    rng = np.random.default_rng(n + bits)
    if bits == 64:
        keys = rng.integers(0, 2 ** 63, size=n, dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, size=n, dtype=np.uint64)
    else:
        # Zipf-like skew: many repeated keys exercise stability and hot digits
        base = rng.integers(0, 1 << bits, size=max(1, n // 7), dtype=np.uint64)
        keys = base[rng.zipf(1.3, size=n) % len(base)] if n else np.zeros(0, np.uint64)
    vals = np.arange(n, dtype=np.float32)  # input positions (exact below 2^24)
    dk = _dev(torch, keys.astype(np.uint64))
    dv = _dev(torch, vals)
    sptc.check(sptc.lib().sptc_sort_pairs(ctypes.c_void_p(dk.data_ptr() if n else 0),
                                          ctypes.c_void_p(dv.data_ptr() if n else 0), n, bits, rb,
                                          _stream(torch)))
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(_host_u64(dk), keys[order])
    np.testing.assert_array_equal(dv.cpu().numpy(), vals[order])


# ---- scan ------------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 1_000_000, 5_000_001])
def test_exclusive_scan(gpu, n):
    import torch
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, size=n, dtype=np.uint64)
    d_in = _dev(torch, a)
    d_out = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    sptc.check(sptc.lib().sptc_exclusive_scan_u64(ctypes.c_void_p(d_in.data_ptr() if n else 0),
                                                  ctypes.c_void_p(d_out.data_ptr()), n,
                                                  _stream(torch)))
    want = np.concatenate([[0], np.cumsum(a)]).astype(np.uint64)
    np.testing.assert_array_equal(_host_u64(d_out), want)


# ---- (c) index match -------------------------------------------------------

@pytest.mark.parametrize("dims,cx,cy", [((200, 150, 120), [2], [0]),           # direct table
                                        ((300, 2000, 4000), [1, 2], [1, 2]),   # search (23 bits)
                                        ((50, 40, 30), [0, 2], [0, 2])])
def test_index_match_matches_oracle(gpu, dims, cx, cy):
    import torch
    rng = np.random.default_rng(sum(dims))
    X = random_coo(rng, dims, 200_000)
    Y = random_coo(rng, dims if cx == cy else (dims[2], dims[1], dims[0]), 150_000)
    cw = orc.common_cwidth(X.dims, Y.dims, cx, cy)
    xl = orc.layout(X.dims, cx, 1, cw)
    yl = orc.layout(Y.dims, cy, 2, cw)
    xk, _ = orc.sort_pairs(orc.linearize(X.idx, xl[0], xl[1]), X.val)
    yk, _ = orc.sort_pairs(orc.linearize(Y.idx, yl[0], yl[1]), Y.val)
    mpos, mcnt, prefix, ir = orc.index_match(xk, yk, xl[4], yl[3])
    dx, dy = _dev(torch, xk), _dev(torch, yk)
    nx, ny = len(xk), len(yk)
    gp = torch.empty(nx, dtype=torch.int64, device="cuda")
    gc = torch.empty(nx, dtype=torch.int64, device="cuda")
    gpre = torch.empty(nx + 1, dtype=torch.int64, device="cuda")
    gir = ctypes.c_uint64()
    sptc.check(sptc.lib().sptc_index_match(
        ctypes.c_void_p(dx.data_ptr()), nx, ctypes.c_void_p(dy.data_ptr()), ny, xl[4], yl[3],
        ctypes.c_void_p(gp.data_ptr()), ctypes.c_void_p(gc.data_ptr()),
        ctypes.c_void_p(gpre.data_ptr()), ctypes.byref(gir), _stream(torch)))
    assert gir.value == ir
    np.testing.assert_array_equal(_host_u64(gc), mcnt)
    np.testing.assert_array_equal(_host_u64(gpre), prefix)
    np.testing.assert_array_equal(_host_u64(gp), mpos)


# ---- full SpTC -------------------------------------------------------------

def _gpu_coords(res):
    idx, v = res.decode()
    return [a.astype(np.int64) for a in idx], v


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_contract_matches_golden(gpu, case):
    res = sptc.SpTC(case["X"], case["Y"], case["cx"], case["cy"])
    coords, v = _gpu_coords(res)
    assert_same_coo(coords, v, case["zcoords"], case["zval"], rtol=REL_TOL)


_compare_with_oracle = compare_with_oracle


@pytest.mark.parametrize("seed", range(6))
def test_contract_random_matches_oracle(gpu, seed):
    rng = np.random.default_rng(1000 + seed)
    xo = int(rng.integers(2, 5))
    nc = int(rng.integers(1, xo))
    yo = int(rng.integers(max(nc, 2), 5))
    cdims = [int(x) for x in rng.integers(2, 60, size=nc)]
    xdims = [int(x) for x in rng.integers(2, 300, size=xo)]
    ydims = [int(x) for x in rng.integers(2, 300, size=yo)]
    cx = [int(x) for x in rng.permutation(xo)[:nc]]
    cy = [int(x) for x in rng.permutation(yo)[:nc]]
    for a, b, d in zip(cx, cy, cdims):
        xdims[a] = d
        ydims[b] = d
    X = random_coo(rng, xdims, 20_000, positive=bool(seed % 2))
    Y = random_coo(rng, ydims, 15_000, positive=bool(seed % 2))
    res = sptc.SpTC(X, Y, cx, cy)
    r, _ = _compare_with_oracle(X, Y, cx, cy, res)
    assert res.ir_nnz == r["ir"]


def test_chunked_equals_unchunked(gpu):
    rng = np.random.default_rng(7)
    X = random_coo(rng, (60, 50, 40), 60_000)
    full = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1)
    assert full.info.nchunks == 1
    # room for the output plus about a third of the products' ESC buffers:
    # forces several accumulation chunks
    small = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1,
                      mem_budget_bytes=full.nnz * 16 + full.ir_nnz * 8 + (128 << 20))
    assert small.info.nchunks > 1
    a = full.to_host()
    b = small.to_host()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    cnt = sptc.SpTC(X, X, [2], [2], count_only=True)
    assert cnt.nnz == full.nnz and cnt.ir_nnz == full.ir_nnz
    # narrow column space but 20 owners per row: the planner keeps ESC
    auto = sptc.SpTC(X, X, [2], [2])
    assert auto.info.accumulator == 0
    for x, y in zip(a, auto.to_host()):
        np.testing.assert_array_equal(x, y)


def test_automatic_accumulator_rule(gpu):
    """The planner picks dense for few owners per row over long Y segments
    and keeps ESC above 5 owners per row; both give the same bits."""
    rng = np.random.default_rng(21)
    X, Y = dense_friendly(rng)
    auto = sptc.SpTC(X, Y, [1], [0])
    assert auto.info.accumulator == 1
    esc = sptc.SpTC(X, Y, [1], [0], dense_accumulator=-1)
    for x, y in zip(auto.to_host(), esc.to_host()):
        np.testing.assert_array_equal(x, y)
    _compare_with_oracle(X, Y, [1], [0], auto)
    X8, _ = dense_friendly(rng, owners=8)
    assert sptc.SpTC(X8, Y, [1], [0]).info.accumulator == 0


def test_verbose_progress_lines(gpu, capfd):
    """verbose: the accumulation announces its chunks and reports the last
    one (products and outputs so far) on stderr."""
    rng = np.random.default_rng(7)
    X = random_coo(rng, (60, 50, 40), 60_000)
    full = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1)
    capfd.readouterr()
    res = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1, verbose=True,
                    mem_budget_bytes=full.nnz * 16 + full.ir_nnz * 8 + (128 << 20))
    err = capfd.readouterr().err
    n = res.info.nchunks
    assert n > 1
    assert f"[sptc] accumulate: {n} chunks, {res.ir_nnz} products" in err
    assert f"[sptc] chunk {n}/{n}: {res.ir_nnz} of {res.ir_nnz} products, {res.nnz} outputs" in err


def test_plain_small_config(gpu):
    """The 2000x2000x500, 2e5-nnz self-contraction over mode 2: global
    invariants on the whole output plus oracle parity on sampled rows."""
    X = PLAIN_SMALL.make()
    res = sptc.SpTC(X, X, [2], [2])
    # ir = sum over contract index of count^2
    cnt = np.bincount(X.idx[2].astype(np.int64), minlength=X.dims[2]).astype(np.float64)
    assert res.ir_nnz == int((cnt ** 2).sum())
    # value checksum: sum of all outputs = sum_c (sum of X values with c)^2
    s = np.bincount(X.idx[2].astype(np.int64), weights=X.val.astype(np.float64),
                    minlength=X.dims[2])
    assert abs(res.info.value_sum - (s ** 2).sum()) <= 1e-9 * (s ** 2).sum()
    xf, yf, v = res.to_host()
    key = (xf.astype(np.uint64) << np.uint64(res.info.y_nfbits)) | yf
    assert np.all(np.diff(key.astype(np.uint64)) > 0)  # strictly ascending: unique, sorted
    assert abs(v.astype(np.float64).sum() - (s ** 2).sum()) <= 1e-5 * (s ** 2).sum()
    rk = res.row_keys()
    pick = np.unique(np.concatenate([rk[:40], rk[-10:], rk[:: max(1, len(rk) // 50)]]))
    _compare_with_oracle(X, X, [2], [2], res, x_rows=pick)


def test_errors_are_status_codes(gpu):
    rng = np.random.default_rng(3)
    X = random_coo(rng, (10, 10, 10), 50)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(X, X, [5], [2])
    assert e.value.status == sptc.ERR_BAD_ARGS
    big = COOTensor((2 ** 31, 2 ** 31, 2 ** 31), [a for a in X.idx], X.val)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(big, big, [0], [0])
    assert e.value.status == sptc.ERR_KEY_OVERFLOW
    Xb = random_coo(rng, (100, 100, 5), 20_000)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(Xb, Xb, [2], [2], mem_budget_bytes=1 << 20)
    assert e.value.status == sptc.ERR_OOM_BUDGET


def test_empty_and_disjoint(gpu):
    X = COOTensor((5, 6), [np.array([0, 1], np.uint32), np.array([0, 1], np.uint32)],
                  np.array([1.0, 2.0], np.float32))
    Y = COOTensor((6, 4), [np.array([2, 3], np.uint32), np.array([0, 1], np.uint32)],
                  np.array([1.0, 2.0], np.float32))
    res = sptc.SpTC(X, Y, [1], [0])
    assert res.nnz == 0 and res.ir_nnz == 0
    E = COOTensor((5, 6), [np.zeros(0, np.uint32), np.zeros(0, np.uint32)],
                  np.zeros(0, np.float32))
    res = sptc.SpTC(E, Y, [1], [0])
    assert res.nnz == 0


# ---- dense per-row accumulator ----------------------------------------------

SMALL_COL = [c for c in FIXTURES if c["name"] in ("self_o3_c2", "self_o4_c23", "spgemm_o2",
                                                  "pow2_dims", "zipf_heavy", "dim1_mode",
                                                  "mixed_order_o4_o3", "outer_nc0",
                                                  "scalar_full", "xy_o3_c21_c02")]


@pytest.mark.parametrize("case", SMALL_COL, ids=[c["name"] for c in SMALL_COL])
def test_dense_matches_golden(gpu, case):
    res = sptc.SpTC(case["X"], case["Y"], case["cx"], case["cy"], dense_accumulator=1)
    assert res.info.accumulator == 1
    coords, v = _gpu_coords(res)
    assert_same_coo(coords, v, case["zcoords"], case["zval"], rtol=REL_TOL)


@pytest.mark.parametrize("seed", range(4))
def test_dense_equals_esc_bitwise(gpu, seed):
    """Both accumulators sum each output's products in the same order with the
    same compensated FP32 steps: identical indices and identical bits."""
    rng = np.random.default_rng(500 + seed)
    X = random_coo(rng, (400, 300, 50), 80_000, positive=bool(seed % 2))
    Y = random_coo(rng, (50, 60, 90), 60_000, positive=bool(seed % 2))
    d = sptc.SpTC(X, Y, [2], [0], dense_accumulator=1)
    e = sptc.SpTC(X, Y, [2], [0], dense_accumulator=-1)
    assert d.info.accumulator == 1 and e.info.accumulator == 0
    for a, b in zip(d.to_host(), e.to_host()):
        np.testing.assert_array_equal(a, b)
    _compare_with_oracle(X, Y, [2], [0], d)


def test_dense_above_an_lds_row_is_tiled(gpu):
    """A column space wider than one LDS row (300x200 Y free keys > 2^13):
    -d 1 runs the tiled dense accumulator; a count-only run cannot."""
    rng = np.random.default_rng(9)
    X = random_coo(rng, (300, 200, 40), 30_000)
    res = sptc.SpTC(X, X, [2], [2], dense_accumulator=1)
    assert res.info.accumulator == 2
    _compare_with_oracle(X, X, [2], [2], res)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(X, X, [2], [2], dense_accumulator=1, count_only=True)
    assert e.value.status == sptc.ERR_BAD_ARGS


def test_duplicate_coordinates_sum_like_the_oracle(gpu):
    """The .tns reader keeps duplicate entries (io_manager.tpp:210-240); ESC
    sums their products like any equal keys; the LDS dense row refuses them
    (its ordered adds assume distinct columns per owner), so -d 1 runs the
    tiled dense accumulator, which adds them like the reference's atomics."""
    rng = np.random.default_rng(10)
    X = random_coo(rng, (30, 20, 10), 1500)
    dup = COOTensor(X.dims, [np.concatenate([a, a[:200]]) for a in X.idx],
                    np.concatenate([X.val, X.val[:200] * 0.5]))
    res = sptc.SpTC(dup, dup, [2], [2])
    assert res.info.accumulator == 0
    _compare_with_oracle(dup, dup, [2], [2], res)
    d = sptc.SpTC(dup, dup, [2], [2], dense_accumulator=1)
    assert d.info.accumulator == 2
    _compare_with_oracle(dup, dup, [2], [2], d)


def test_row_segments_big_and_small_rows(gpu):
    """Zipf rows: a few X rows with far more than half a sort tile of
    products (multi-tile column-bit passes) among many small ones (grouped
    and sorted in LDS), signed values, against the oracle."""
    from gsparc_amd.coo import UNIFORM, ZIPF, VAL_UNIFORM01, generate
    X = generate(0x5EED00AA, (300, 200, 48), (ZIPF, ZIPF, UNIFORM), (1.3, 1.3, 1.0), 40_000,
                 VAL_UNIFORM01)
    rng = np.random.default_rng(12)
    X = COOTensor(X.dims, X.idx, X.val * np.where(rng.random(X.nnz) < 0.5, -1, 1).astype(
        np.float32))
    res = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1)
    prod_per_c = np.bincount(X.idx[2].astype(np.int64), minlength=48)
    rk = X.idx[0].astype(np.int64) * 200 + X.idx[1]
    per_row = np.bincount(np.unique(rk, return_inverse=True)[1],
                          weights=prod_per_c[X.idx[2].astype(np.int64)])
    assert per_row.max() > 3 * 6144 and np.median(per_row) < 3072  # both kinds present
    _compare_with_oracle(X, X, [2], [2], res)
    # count-only runs sort the columns alone: same count, same value sum
    c = sptc.SpTC(X, X, [2], [2], dense_accumulator=-1, count_only=True)
    assert c.nnz == res.nnz and c.ir_nnz == res.ir_nnz
    # (FP64 sums added by workgroups in any order)
    assert np.isclose(c.info.value_sum, res.info.value_sum, rtol=1e-12, atol=0)


@pytest.mark.parametrize("ycols", [1, 3, 40])
def test_long_runs_and_row_starts_across_wave_tiles(gpu, ycols):
    """Reduce-tile edge cases, against the oracle: runs of hundreds to
    thousands of products (Y has only `ycols` free columns) that cross lanes
    and 512-product wave tiles, rows whose products end at arbitrary offsets
    (sizes from 1 to 8000 products, big rows and small-row groups), rows with
    no products (X nonzeros whose contracted index Y lacks), signed values."""
    rng = np.random.default_rng(100 + ycols)
    C = 3000
    rows, xs = [], []
    for r, n in enumerate([1, 2, 7, 64, 511, 512, 513, 900, 1500, 2600, 4000, 9000, 20, 3, 70]):
        c = rng.choice(C, size=min(n, C), replace=False)
        rows.append(np.full(len(c), r, np.uint32))
        xs.append(c.astype(np.uint32))
    xi, xc = np.concatenate(rows), np.concatenate(xs)
    xv = (rng.random(len(xi)) - 0.5).astype(np.float32)
    X = COOTensor((len(rows), C), [xi, xc], xv)
    # Y: contracted index c in [0, 2700) only (X nonzeros at c >= 2700 give
    # no products), free column j in [0, ycols)
    yc, yj = [], []
    for c in range(2700):
        k = int(rng.integers(1, ycols + 1)) if ycols > 1 else 1
        for j in rng.choice(ycols, size=k, replace=False):
            yc.append(c)
            yj.append(j)
    yc, yj = np.asarray(yc, np.uint32), np.asarray(yj, np.uint32)
    yv = (rng.random(len(yc)) - 0.5).astype(np.float32)
    Y = COOTensor((C, ycols), [yc, yj], yv)
    res = sptc.SpTC(X, Y, [1], [0], dense_accumulator=-1)
    _compare_with_oracle(X, Y, [1], [0], res)
    c = sptc.SpTC(X, Y, [1], [0], dense_accumulator=-1, count_only=True)
    assert c.nnz == res.nnz
    assert np.isclose(c.info.value_sum, res.info.value_sum, rtol=1e-12, atol=1e-12)


def bucket_case(name):
    """Big rows (tens of thousands of products each) over wide column spaces:
    the shapes of the high-digit bucket reduce (SPTC_BIG_TAIL=buckets).
    zipf15: Y columns drawn Zipf, so
    a bucket holds one column from many owners, often several inside one
    64-product window (signed values); wide20: 20 column bits, buckets of
    1024 columns; low11: 11 column bits, buckets of 2 columns."""
    from gsparc_amd.coo import UNIFORM, ZIPF, VAL_UNIFORM01, generate
    if name == "zipf15":
        X = generate(0xB0C1, (24, 3000), (UNIFORM, UNIFORM), (1.0, 1.0), 24 * 500, VAL_UNIFORM01)
        Y = generate(0xB0C2, (3000, 300_000), (UNIFORM, ZIPF), (1.0, 1.1), 3000 * 40,
                     VAL_UNIFORM01)
        rng = np.random.default_rng(17)
        Y = COOTensor(Y.dims, Y.idx, Y.val * np.where(rng.random(Y.nnz) < 0.5, -1, 1).astype(
            np.float32))
    elif name == "wide20":
        X = generate(0xB0C3, (16, 3000), (UNIFORM, UNIFORM), (1.0, 1.0), 16 * 300, VAL_UNIFORM01)
        Y = generate(0xB0C4, (3000, 1 << 20), (UNIFORM, UNIFORM), (1.0, 1.0), 3000 * 250,
                     VAL_UNIFORM01)
    else:
        X = generate(0xB0C5, (24, 3000), (UNIFORM, UNIFORM), (1.0, 1.0), 24 * 500, VAL_UNIFORM01)
        Y = generate(0xB0C6, (3000, 2000), (UNIFORM, UNIFORM), (1.0, 1.0), 3000 * 40,
                     VAL_UNIFORM01)
    return X, Y


BUCKET_CASES = ["zipf15", "wide20", "low11"]


@pytest.mark.parametrize("name", BUCKET_CASES)
def test_big_rows_wide_columns_against_oracle(gpu, name):
    """Big rows over 11 to 20 column bits (two column passes, or with
    SPTC_BIG_TAIL=buckets one pass and buckets of 2 to 1024 columns),
    against the oracle, count-only included."""
    X, Y = bucket_case(name)
    res = sptc.SpTC(X, Y, [1], [0], dense_accumulator=-1)
    assert res.info.accumulator == 0
    bits = int(len(np.unique(Y.idx[1])) - 1).bit_length()  # column ranks
    assert bits == {"zipf15": 15, "wide20": 20, "low11": 11}[name], bits
    _compare_with_oracle(X, Y, [1], [0], res)
    c = sptc.SpTC(X, Y, [1], [0], dense_accumulator=-1, count_only=True)
    assert c.nnz == res.nnz


def test_index_range_check(gpu):
    """check_indices: an index at or above its mode's dim is an error; by
    default it is masked to the mode's width, as the reference's pdep
    linearization does."""
    X = COOTensor((4, 5, 6), [np.array([0, 3], np.uint32), np.array([1, 2], np.uint32),
                              np.array([5, 6], np.uint32)], np.array([1.0, 2.0], np.float32))
    with pytest.raises(sptc.SpTCError) as e:
        sptc.SpTC(X, X, [2], [2], check_indices=True)
    assert e.value.status == sptc.ERR_BAD_ARGS and "mode 2" in str(e.value)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.Plan(X, X, [2], [2], check_indices=True)
    assert e.value.status == sptc.ERR_BAD_ARGS
    sptc.SpTC(X, X, [2], [2])  # masked, no error
    ok = COOTensor((4, 5, 7), X.idx, X.val)
    assert sptc.SpTC(ok, ok, [2], [2], check_indices=True).nnz == 2
