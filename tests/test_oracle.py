"""Pins the CPU oracle (oracle/sptc_oracle.c) against what the mathematics and
the hardware fix: dense tensordot outputs (golden fixtures), BMI2 pdep known
answers, closed-form key layouts, stable sorting, brute-force matching, and
partitioning invariants.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as orc
from helpers import (GOLDEN, REL_TOL, assert_same_coo, decode_x, decode_y, dense,
                     load_fixtures, max_rel_err, random_coo)

FIXTURES = load_fixtures()


# ---- layout (tensor_manager.tpp:63-186) ------------------------------------

def test_layout_closed_form_example():
    # dims 2000x2000x500, contract mode 2: widths 11, 11, 9
    sh, wd, fp, nf, nc = orc.layout((2000, 2000, 500), [2], 1)
    assert wd == [11, 11, 9] and fp == [0, 1] and (nf, nc) == (22, 9)
    assert sh == [9 + 11, 9, 0]  # X: free (last free mode at bit 0) above contract
    sh, wd, fp, nf, nc = orc.layout((2000, 2000, 500), [2], 2)
    assert sh == [11, 0, 22]  # Y: contract above free


def test_layout_mode_bits():
    for d, b in [(1, 0), (2, 1), (3, 2), (4, 2), (5, 3), (2048, 11), (2049, 12), (17, 5),
                 (14036, 14), (2 ** 32, 32)]:
        assert orc.mode_bits(d) == b
        assert (d - 1) < (1 << b) or d == 1


@pytest.mark.parametrize("seed", range(20))
def test_layout_fields_tile_the_key(seed):
    rng = np.random.default_rng(seed)
    order = int(rng.integers(1, 7))
    dims = [int(x) for x in rng.integers(1, 5000, size=order)]
    nc = int(rng.integers(0, order + 1))
    cm = [int(x) for x in rng.permutation(order)[:nc]]
    for role in (1, 2):
        sh, wd, fp, nf, ncb = orc.layout(dims, cm, role)
        masks = [((1 << wd[m]) - 1) << sh[m] for m in range(order)]
        union = 0
        for m in range(order):
            assert union & masks[m] == 0  # fields disjoint
            union |= masks[m]
        assert union == (1 << (nf + ncb)) - 1  # and contiguous from bit 0
        cmask = sum(masks[m] for m in cm)
        if role == 1:
            assert cmask == (1 << ncb) - 1  # contract field at the bottom
        else:
            assert cmask == ((1 << ncb) - 1) << nf
        # contract modes: caller order, last one lowest
        for a, b in zip(cm, cm[1:]):
            assert sh[a] > sh[b]
        for a, b in zip(fp, fp[1:]):
            assert sh[a] > sh[b]


# ---- pdep / linearize (bitops.hpp:9-13, tensor_manager.tpp:212-298) --------

def test_pdep_matches_bmi2_known_answers():
    with open(os.path.join(GOLDEN, "pdep_kat.json")) as f:
        cases = json.load(f)["cases"]
    assert len(cases) >= 100
    for x, m, want in cases:
        assert orc.pdep(int(x), int(m)) == int(want)


@pytest.mark.parametrize("seed", range(6))
def test_linearize_closed_form_and_roundtrip(seed):
    rng = np.random.default_rng(100 + seed)
    order = int(rng.integers(2, 6))
    dims = [int(x) for x in rng.integers(2, 3000, size=order)]
    nc = int(rng.integers(1, order))
    cm = [int(x) for x in rng.permutation(order)[:nc]]
    t = random_coo(rng, dims, 500)
    for role in (1, 2):
        sh, wd, fp, nf, ncb = orc.layout(dims, cm, role)
        keys = orc.linearize(t.idx, sh, wd)
        want = np.zeros(t.nnz, np.uint64)
        for m in range(order):
            want |= t.idx[m].astype(np.uint64) << np.uint64(sh[m])
        np.testing.assert_array_equal(keys, want)
        for m in range(order):
            got = (keys >> np.uint64(sh[m])) & np.uint64((1 << wd[m]) - 1)
            np.testing.assert_array_equal(got, t.idx[m].astype(np.uint64))


# ---- sort (sort.cuh:39-159) ------------------------------------------------

@pytest.mark.parametrize("n,bits", [(0, 10), (1, 5), (1000, 7), (20000, 40), (5000, 64)])
def test_sort_pairs_is_stable(n, bits):
    rng = np.random.default_rng(n + bits)
    hi = 1 << min(bits, 63)
    keys = rng.integers(0, hi, size=n, dtype=np.uint64)
    if bits == 64:
        keys |= rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)
    vals = np.arange(n, dtype=np.float32)  # values record the input position
    k, v = orc.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


# ---- index match (indexMatch.cuh:20-118, 231-287) --------------------------

@pytest.mark.parametrize("seed", range(5))
def test_index_match_bruteforce(seed):
    rng = np.random.default_rng(200 + seed)
    X = random_coo(rng, (20, 15, 12), 400)
    Y = random_coo(rng, (12, 9, 15), 300)
    cx, cy = [2, 1], [0, 2]
    cw = orc.common_cwidth(X.dims, Y.dims, cx, cy)
    xl = orc.layout(X.dims, cx, 1, cw)
    yl = orc.layout(Y.dims, cy, 2, cw)
    xk, _ = orc.sort_pairs(orc.linearize(X.idx, xl[0], xl[1]), X.val)
    yk, _ = orc.sort_pairs(orc.linearize(Y.idx, yl[0], yl[1]), Y.val)
    ncb, ynf = xl[4], yl[3]
    mpos, mcnt, prefix, ir = orc.index_match(xk, yk, ncb, ynf)
    yc = yk >> np.uint64(ynf)
    for i, k in enumerate(xk):
        c = k & np.uint64((1 << ncb) - 1)
        hits = np.nonzero(yc == c)[0]
        assert mcnt[i] == len(hits)
        if len(hits):
            assert mpos[i] == hits[0] and np.all(np.diff(hits) == 1)
    assert prefix[0] == 0 and prefix[-1] == ir
    np.testing.assert_array_equal(np.diff(prefix.astype(np.int64)), mcnt.astype(np.int64))
    # ir = sum_c cX(c) * cY(c) over contract coordinates
    xcnt = {}
    for i in range(X.nnz):
        key = (int(X.idx[2][i]), int(X.idx[1][i]))
        xcnt[key] = xcnt.get(key, 0) + 1
    ycnt = {}
    for j in range(Y.nnz):
        key = (int(Y.idx[0][j]), int(Y.idx[2][j]))
        ycnt[key] = ycnt.get(key, 0) + 1
    assert ir == sum(v * ycnt.get(k, 0) for k, v in xcnt.items())


# ---- full pipeline vs dense tensordot (contraction.cuh:100-141, 364-599) ----

def _oracle_coo(r):
    sh, wd, fp, nf, ncb = r["xl"]
    ysh, ywd, yfp, ynf, _ = r["yl"]
    coords = decode_x(r["xf"], sh, wd, fp, ncb) + decode_y(r["yf"], ysh, ywd, yfp)
    return coords


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_oracle_pipeline_matches_golden(case):
    r = orc.pipeline(case["X"], case["Y"], case["cx"], case["cy"])
    coords = _oracle_coo(r)
    assert_same_coo(coords, r["v64"], case["zcoords"], case["zval"], rtol=1e-12)
    assert max_rel_err(r["v32"], case["zval"]) <= REL_TOL


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_oracle_join_matches_golden(case):
    X, Y = case["X"], case["Y"]
    xf, yf, v = orc.join_fp64(X, Y, case["cx"], case["cy"])
    zx = [d for m, d in enumerate(X.dims) if m not in case["cx"]]
    zy = [d for m, d in enumerate(Y.dims) if m not in case["cy"]]
    coords = []
    if zx:
        coords += [a.astype(np.int64) for a in np.unravel_index(xf.astype(np.int64), zx)]
    if zy:
        coords += [a.astype(np.int64) for a in np.unravel_index(yf.astype(np.int64), zy)]
    assert_same_coo(coords, v, case["zcoords"], case["zval"], rtol=1e-12)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_pipeline_random_vs_tensordot(seed):
    rng = np.random.default_rng(300 + seed)
    xo = int(rng.integers(2, 5))
    nc = int(rng.integers(1, xo))
    yo = int(rng.integers(nc, 5))
    cdims = [int(x) for x in rng.integers(2, 7, size=nc)]
    xdims = [int(x) for x in rng.integers(2, 7, size=xo)]
    ydims = [int(x) for x in rng.integers(2, 7, size=yo)]
    cx = [int(x) for x in rng.permutation(xo)[:nc]]
    cy = [int(x) for x in rng.permutation(yo)[:nc]]
    for a, b, d in zip(cx, cy, cdims):
        xdims[a] = d
        ydims[b] = d
    X = random_coo(rng, xdims, int(np.prod(xdims)) // 3 + 1, positive=False)
    Y = random_coo(rng, ydims, int(np.prod(ydims)) // 3 + 1, positive=False)
    r = orc.pipeline(X, Y, cx, cy)
    Z = np.tensordot(dense(X), dense(Y), axes=(cx, cy))
    coords = _oracle_coo(r)
    # every product-bearing coordinate appears (ESC keeps structural zeros)
    Zd = np.zeros(Z.shape)
    Zd[tuple(coords)] = r["v64"]
    np.testing.assert_allclose(Zd, Z, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(Z) <= len(r["v64"])


def test_dense_nnz_drops_exact_zeros():
    assert orc.dense_nnz(np.array([0.0, 1.0, -0.0, 2.5, 0.0], np.float32)) == 2


# ---- partitioning (tensor_manager.tpp:633-749, contraction.cuh:143-245) ----

@pytest.mark.parametrize("prtn", [1, 2, 3, 7, 50])
def test_partition_never_splits_a_row(prtn):
    rng = np.random.default_rng(prtn)
    keys = np.sort(rng.integers(0, 1 << 20, size=3000, dtype=np.uint64))
    shift = 8
    b = orc.partition(keys, prtn, shift)
    assert b[0] == 0 and b[-1] == len(keys)
    assert np.all(np.diff(b.astype(np.int64)) > 0)
    for e in b[1:-1]:
        assert (keys[e] >> np.uint64(shift)) != (keys[e - 1] >> np.uint64(shift))


@pytest.mark.parametrize("avail", [10 ** 9, 200000, 30000])
def test_dynamic_partition_ranges_fit(avail):
    case = FIXTURES[-1]
    r = orc.pipeline(case["X"], case["X"], [1], [1])
    xk, prefix, ir = r["xkeys"], r["prefix"], r["ir"]
    ncb = r["xl"][4]
    ends, mx = orc.dynamic_partition(xk, ir, prefix, ncb, avail)
    assert ends is not None and ends[-1] == len(xk) - 1
    start = 0
    for e in ends:
        e = int(e)
        if e < len(xk) - 1:
            assert (xk[e] >> np.uint64(ncb)) != (xk[e + 1] >> np.uint64(ncb))
        rng_ir = int(prefix[e + 1] - prefix[start])
        if start != 0 or e != len(xk) - 1:
            last_row = (xk[start] >> np.uint64(ncb)) == (xk[len(xk) - 1] >> np.uint64(ncb))
            if not last_row:
                assert 2 * (rng_ir * 24 + 8) <= avail
        assert rng_ir <= mx
        start = e + 1
