"""GPU parity of the five stated workloads at their stated sizes.

Each full output is pinned by closed forms that need no products (the
product count ir = sum of count(c)^2-type sums, the value checksum
sum_x v_x * S_Y(c(x))), by a device pass over every emitted entry (FP64 sum
of the values, strict (row, column) order), and by sampled X rows compared
with the CPU oracle: indices bit-exact, values within REL_TOL of FP64.
"""
import time

import numpy as np
import pytest

from gsparc_amd import sptc
from gsparc_amd.coo import COOTensor
from gsparc_amd.workloads import (NELL2, NELL2_CAP_BYTES, NIPS, batch_instance,
                                  serving_large)
from helpers import (U32, closed_forms, compare_with_oracle, pick_rows, x_free_keys)

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[configs] {msg}", flush=True)


def _check_full(res, X, Y, cx, cy, ir=None, vsum=None):
    """Closed forms and the device verification pass of a full result."""
    if ir is None:
        ir, vsum = closed_forms(X, Y, cx, cy)
    assert res.ir_nnz == ir
    if ir:
        assert abs(res.info.value_sum - vsum) <= 2 * U32 * vsum
    emitted, bad = res.verify()
    assert bad == 0, f"{bad} entries out of (row, column) order"
    assert abs(emitted - vsum) <= 6 * U32 * vsum + 1e-300
    return ir, vsum


def _sub(X, mask):
    return COOTensor(X.dims, [a[mask] for a in X.idx], X.val[mask])


# ---- NIPS-shaped: order 4, 3.1e6 nnz, contract {2,3} ------------------------

def test_nips_full_size(gpu):
    X = NIPS.make()
    cx = cy = list(NIPS.cx)
    t0 = time.time()
    res = sptc.SpTC(X, X, cx, cy)
    _log(f"nips: ir {res.ir_nnz} out {res.nnz} chunks {res.info.nchunks} "
         f"({time.time() - t0:.1f} s)")
    ir, vsum = _check_full(res, X, X, cx, cy)
    assert ir == 12_990_190_978  # sum_c count(c)^2 of this seeded tensor
    rows = pick_rows(X, X, cx, cy)
    _, err = compare_with_oracle(X, X, cx, cy, res, x_rows=rows)
    _log(f"nips: {len(rows)} sampled rows match the oracle, max rel err {err:.2e}")


# ---- NELL-2-shaped: 7.69e7 nnz, contract mode 2, 64 GB cap -----------------

@pytest.fixture(scope="module")
def nell2():
    t0 = time.time()
    X = NELL2.make()
    _log(f"nell2 generated in {time.time() - t0:.1f} s")
    return X


def test_nell2_full_match_and_partition(gpu, nell2):
    """Index match over the whole 7.69e7-nnz operand and the chunk plan of
    the full 3e13-product contraction under the 64 GB cap."""
    X = nell2
    cx = cy = list(NELL2.cx)
    p = sptc.Plan(X, X, cx, cy, mem_budget_bytes=NELL2_CAP_BYTES)
    rows, prods, ir = p.partitions(budget_bytes=NELL2_CAP_BYTES)
    want_ir, _ = closed_forms(X, X, cx, cy)
    assert ir == want_ir
    assert prods[0] == 0 and prods[-1] == ir
    assert np.all(np.diff(rows.astype(np.int64)) > 0)
    assert np.all(np.diff(prods.astype(np.int64)) >= 0)
    assert rows[-1] == len(np.unique(x_free_keys(X, X, cx, cy)))
    _log(f"nell2: ir {ir:.4e} in {len(rows) - 1} chunks under 64 GB")


def test_nell2_row_sample_count_only(gpu, nell2):
    """Count-only contraction of every 40th X row against the whole tensor,
    under the 64 GB cap (several chunks), against the closed forms."""
    X = nell2
    cx = cy = list(NELL2.cx)
    keys = x_free_keys(X, X, cx, cy)
    u, inv = np.unique(keys, return_inverse=True)
    Xs = _sub(X, (inv % 40) == 0)
    ir, vsum = closed_forms(Xs, X, cx, cy)
    res = sptc.SpTC(Xs, X, cx, cy, count_only=True, mem_budget_bytes=NELL2_CAP_BYTES)
    assert res.ir_nnz == ir
    assert abs(res.info.value_sum - vsum) <= 2 * U32 * vsum
    assert res.info.nchunks > 1 and res.nnz > 0
    _log(f"nell2 sample: {Xs.nnz} nnz, ir {ir:.3e}, out {res.nnz}, {res.info.nchunks} chunks")


def test_nell2_rows_full_output(gpu, nell2):
    """Full output of a few X rows against the whole tensor vs the oracle."""
    X = nell2
    cx = cy = list(NELL2.cx)
    keys = x_free_keys(X, X, cx, cy)
    u = np.unique(keys)
    pick = np.isin(keys, u[[0, 1, 7, len(u) // 3, len(u) // 2, len(u) - 1]])
    Xs = _sub(X, pick)
    res = sptc.SpTC(Xs, X, cx, cy, mem_budget_bytes=NELL2_CAP_BYTES)
    _check_full(res, Xs, X, cx, cy)
    _, err = compare_with_oracle(Xs, X, cx, cy, res)
    _log(f"nell2 rows: {Xs.nnz} nnz, out {res.nnz}, max rel err {err:.2e}")


# ---- the 256-instance batch, instances as drawn -----------------------------

BATCH_FULL_IR = 1 << 28  # instances up to this many products run batched, in full


@pytest.fixture(scope="module")
def batch_as_stated():
    out = []
    for i in range(256):
        s = batch_instance(i)
        X = s.make()
        ir, vsum = closed_forms(X, X, list(s.cx), list(s.cy))
        out.append((i, s, X, ir, vsum))
    return out


def test_batch_as_stated_batched(gpu, batch_as_stated):
    """The instances of the stated distribution with at most 2^28 products,
    as drawn (no nnz scaling), in one batch: per-instance closed forms and
    order checks, sampled rows of every 12th instance vs the oracle."""
    items = [t for t in batch_as_stated if t[3] <= BATCH_FULL_IR]
    b = sptc.Batch([(X, X, list(s.cx), list(s.cy)) for _, s, X, _, _ in items])
    info = b.info
    _log(f"batch: {len(items)} instances, ir {info['ir_nnz']:.3e}, out {info['out_nnz']:.3e}")
    worst = 0.0
    for k, (i, s, X, ir, vsum) in enumerate(items):
        r = b.result(k)
        _check_full(r, X, X, list(s.cx), list(s.cy), ir, vsum)
        if k % 12 == 0 and ir:
            _, err = compare_with_oracle(X, X, list(s.cx), list(s.cy), r,
                                         x_rows=pick_rows(X, X, list(s.cx), list(s.cy),
                                                          total_cap=3_000_000))
            worst = max(worst, err)
        del r
    _log(f"batch: closed forms hold for all; oracle rows max rel err {worst:.2e}")


def test_batch_as_stated_large_instances(gpu, batch_as_stated):
    """The instances above 2^28 products (65 of 256; one alone has 1.2e11)
    run as plain calls: count-only above 2^31 products (their outputs would
    exceed device memory), in full below, against the closed forms."""
    big = [t for t in batch_as_stated if t[3] > BATCH_FULL_IR]
    total = 0
    for i, s, X, ir, vsum in big:
        full = ir <= (1 << 31)
        res = sptc.SpTC(X, X, list(s.cx), list(s.cy), count_only=not full)
        assert res.ir_nnz == ir, i
        assert abs(res.info.value_sum - vsum) <= 2 * U32 * vsum, i
        if full:
            _check_full(res, X, X, list(s.cx), list(s.cy), ir, vsum)
        total += ir
        del res
    _log(f"batch large: {len(big)} instances, {total:.3e} products, closed forms hold")


# ---- the serving mix ------------------------------------------------------------

def test_serving_mix_against_oracle(gpu, batch_as_stated):
    """A request mix of the serving distribution -- 90 % instances of the
    batched distribution, 10 % large (order 3, dims ~1e4, 1e7 nnz) -- served
    together under the 96 GB budget; every full result is checked against
    the closed forms and the order pass, some against oracle rows, and the
    count-only (large) ones against the closed forms."""
    rng = np.random.default_rng(4)
    small = [batch_as_stated[int(i)] for i in rng.choice(256, size=18, replace=False)]
    large = [serving_large(k) for k in range(2)]
    srv = sptc.Server(budget_bytes=96 << 30)
    reqs = []
    for i, s, X, ir, vsum in small:
        tid = srv.add_tensor(X)
        reqs.append((X, list(s.cx), ir, vsum, srv.submit(tid, tid, list(s.cx), list(s.cy))))
    for k, s in enumerate(large):
        X = s.make()
        ir, vsum = closed_forms(X, X, list(s.cx), list(s.cy))
        tid = srv.add_tensor(X)
        # interleave the large requests with the small ones
        reqs.insert(6 * k + 3, (X, list(s.cx), ir, vsum,
                                srv.submit(tid, tid, list(s.cx), list(s.cy))))
    checked = 0
    for k, (X, cm, ir, vsum, rid) in enumerate(reqs):
        r, info = srv.wait(rid)
        assert info["status"] == 0
        assert r.ir_nnz == ir
        assert abs(r.info.value_sum - vsum) <= 2 * U32 * max(vsum, 1e-300)
        if not info["count_only"]:
            _check_full(r, X, X, cm, cm, ir, vsum)
            if k % 4 == 0 and ir:
                compare_with_oracle(X, X, cm, cm, r,
                                    x_rows=pick_rows(X, X, cm, cm, total_cap=2_000_000))
                checked += 1
        del r
    st = srv.stats()
    assert st["large_requests"] >= 2
    assert st["peak_in_use_bytes"] <= st["budget_bytes"]
    srv.close()
    _log(f"serving: {len(reqs)} requests, {checked} compared with the oracle, "
         f"{st['batches']} batches, {st['large_requests']} served alone")
