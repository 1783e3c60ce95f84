"""Shared helpers for the SpTC tests: fixtures, key decoding, comparisons."""
import json
import os

import numpy as np

import oracle as orc
from gsparc_amd.coo import COOTensor

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Relative-error bound on FP32 output values against the FP64 oracle.
REL_TOL = 1e-5


def load_fixtures():
    with open(os.path.join(GOLDEN, "sptc_fixtures.json")) as f:
        cases = json.load(f)["cases"]
    out = []
    for c in cases:
        def t(d):
            return COOTensor(tuple(d["dims"]), [np.asarray(a, np.uint32) for a in d["idx"]],
                             np.asarray(d["val"], np.float32))
        X = t(c["X"])
        Y = t(c["Y"]) if c["Y"] else X
        Z = c["Z"]
        coords = [np.asarray(a, np.int64) for a in Z["coords"]]
        out.append(dict(name=c["name"], X=X, Y=Y, cx=c["cx"], cy=c["cy"], zcoords=coords,
                        zval=np.asarray(Z["val"], np.float64), zdims=Z["dims"]))
    return out


def decode_x(xf, shift, width, fpos, ncbits):
    """X free key (key >> ncbits) -> list of per-mode indices (free modes
    ascending)."""
    out = []
    for m in fpos:
        s = shift[m] - ncbits
        out.append(((xf.astype(np.uint64) >> np.uint64(s)) &
                    np.uint64((1 << width[m]) - 1)).astype(np.int64))
    return out


def decode_y(yf, shift, width, fpos):
    out = []
    for m in fpos:
        out.append(((yf.astype(np.uint64) >> np.uint64(shift[m])) &
                    np.uint64((1 << width[m]) - 1)).astype(np.int64))
    return out


def lex_order(coords):
    if not coords:
        return np.arange(0)
    return np.lexsort(tuple(reversed(coords)))


def assert_same_coo(coords_a, vals_a, coords_b, vals_b, rtol=REL_TOL):
    """Index sets equal (bit-exact) and values within rtol (relative)."""
    assert len(vals_a) == len(vals_b), (len(vals_a), len(vals_b))
    assert len(coords_a) == len(coords_b)
    oa = lex_order(coords_a) if coords_a else np.arange(len(vals_a))
    ob = lex_order(coords_b) if coords_b else np.arange(len(vals_b))
    for a, b in zip(coords_a, coords_b):
        np.testing.assert_array_equal(np.asarray(a)[oa], np.asarray(b)[ob])
    va = np.asarray(vals_a, np.float64)[oa]
    vb = np.asarray(vals_b, np.float64)[ob]
    if len(va):
        rel = np.abs(va - vb) / np.maximum(np.abs(vb), 1e-300)
        assert rel.max() <= rtol, f"max rel err {rel.max():.3e} > {rtol}"


def max_rel_err(v, ref):
    v = np.asarray(v, np.float64)
    ref = np.asarray(ref, np.float64)
    if not len(v):
        return 0.0
    return float((np.abs(v - ref) / np.maximum(np.abs(ref), 1e-300)).max())


def dense(t):
    D = np.zeros(t.dims, dtype=np.float64)
    D[tuple(a.astype(np.int64) for a in t.idx)] = t.val.astype(np.float64)
    return D


def random_coo(rng, dims, nnz, positive=True):
    space = int(np.prod(dims))
    nnz = min(nnz, space)
    lin = rng.choice(space, size=nnz, replace=False)
    idx = [a.astype(np.uint32) for a in np.unravel_index(lin, dims)]
    val = rng.random(nnz).astype(np.float32) + np.float32(0.05)
    if not positive:
        val = val * np.where(rng.random(nnz) < 0.5, -1, 1).astype(np.float32)
    return COOTensor(tuple(int(d) for d in dims), idx, val)


def abs_values(t):
    return COOTensor(t.dims, t.idx, np.abs(t.val))


def compare_with_oracle(X, Y, cx, cy, res, x_rows=None):
    """Indices bit-exact; values within REL_TOL of the FP64 oracle relative
    to the sum of |products| of each output (equal to the plain relative
    error when values are positive; with signed values a cancelling sum
    makes the plain relative error unbounded)."""
    r = orc.pipeline(X, Y, cx, cy, x_rows=x_rows)
    signed = bool((X.val < 0).any() or (Y.val < 0).any())
    scale = (orc.pipeline(abs_values(X), abs_values(Y), cx, cy, x_rows=x_rows)["v64"]
             if signed else None)
    if x_rows is None:
        xf, yf, v = res.to_host()
    else:
        rk = res.row_keys()
        parts = []
        for key in np.sort(np.asarray(x_rows, np.uint64)):
            j = int(np.searchsorted(rk, key))
            if j < len(rk) and rk[j] == key:
                b, e = res.row_range(j)
                parts.append(res.copy_range(b, e))
        xf = np.concatenate([p[0] for p in parts]) if parts else np.zeros(0, np.uint64)
        yf = np.concatenate([p[1] for p in parts]) if parts else np.zeros(0, np.uint64)
        v = np.concatenate([p[2] for p in parts]) if parts else np.zeros(0, np.float32)
    np.testing.assert_array_equal(xf, r["xf"])
    np.testing.assert_array_equal(yf, r["yf"])
    if signed:
        err = float((np.abs(v.astype(np.float64) - r["v64"]) / scale).max()) if len(v) else 0.0
    else:
        err = max_rel_err(v, r["v64"])
    assert err <= REL_TOL, err
    return r, err


# ---- closed forms of a contraction (exact index arithmetic in numpy) -------

# FP32 unit roundoff.  With positive values an FP64 sum of FP32-rounded
# products is within U32 (relative) of the exact sum, and a compensated FP32
# sum of them within about 3 * U32.
U32 = 2.0 ** -24


def contract_keys(t, cmodes):
    """Row-major linear id of each nonzero's contracted coordinates."""
    k = np.zeros(t.nnz, np.int64)
    for m in cmodes:
        k = k * np.int64(t.dims[m]) + t.idx[m].astype(np.int64)
    return k


def closed_forms(X, Y, cx, cy):
    """(ir, value_sum) of X x_{cx,cy} Y without forming products:
    ir = sum over X nonzeros of the count of Y nonzeros with the same
    contracted coordinates; value_sum = sum_x v_x * S_Y(c(x)) with S_Y(c) the
    sum of Y's values at c (FP64)."""
    kx, ky = contract_keys(X, cx), contract_keys(Y, cy)
    uniq, inv = np.unique(np.concatenate([kx, ky]), return_inverse=True)
    ix, iy = inv[:len(kx)], inv[len(kx):]
    cnt = np.bincount(iy, minlength=len(uniq)).astype(np.int64)
    s = np.bincount(iy, weights=Y.val.astype(np.float64), minlength=len(uniq))
    return int(cnt[ix].sum()), float((X.val.astype(np.float64) * s[ix]).sum())


def x_free_keys(X, Y, cx, cy):
    """Each X nonzero's X free key in the contraction's layout (the row keys
    of the result, and the x_rows argument of the oracle pipeline)."""
    cw = orc.common_cwidth(X.dims, Y.dims, cx, cy)
    sh, wd, fp, nf, ncb = orc.layout(X.dims, cx, 1, cw)
    k = np.zeros(X.nnz, np.uint64)
    for m in fp:
        k |= X.idx[m].astype(np.uint64) << np.uint64(sh[m] - ncb)
    return k


def pick_rows(X, Y, cx, cy, per_row_cap=2_000_000, total_cap=12_000_000, strided=40):
    """X free rows to compare with the oracle: the first and last rows, rows
    taken at a stride, and the heaviest rows under per_row_cap products (the
    multi-tile sort path), keeping the oracle's total work bounded."""
    kx, ky = contract_keys(X, cx), contract_keys(Y, cy)
    uniq, inv = np.unique(np.concatenate([kx, ky]), return_inverse=True)
    cnt = np.bincount(inv[len(kx):], minlength=len(uniq)).astype(np.int64)
    prod = cnt[inv[:len(kx)]]
    rows, rinv = np.unique(x_free_keys(X, Y, cx, cy), return_inverse=True)
    rprod = np.bincount(rinv, weights=prod.astype(np.float64), minlength=len(rows))
    cand = list(range(min(12, len(rows)))) + list(range(max(0, len(rows) - 12), len(rows)))
    cand += list(range(0, len(rows), max(1, len(rows) // strided)))
    heavy = [int(i) for i in np.argsort(-rprod) if rprod[i] <= per_row_cap][:6]
    pick, total = [], 0.0
    for i in heavy + cand:
        if i in pick or rprod[i] > per_row_cap or total + rprod[i] > total_cap:
            continue
        pick.append(i)
        total += rprod[i]
    return rows[np.array(sorted(pick), np.int64)]


def dense_friendly(rng, rows=500, owners=2, cdim=20, ycols=200, per_c=150):
    """Few active X nonzeros per row (`owners`), long Y segments (`per_c`
    entries per contract index) over a narrow column space: the shape the
    automatic rule sends to the dense accumulator."""
    xr = np.repeat(np.arange(rows, dtype=np.uint32), owners)
    xc = rng.integers(0, cdim, size=len(xr)).astype(np.uint32)
    X = COOTensor((rows, cdim), [xr, xc], (rng.random(len(xr)) + 0.1).astype(np.float32))
    yc = np.repeat(np.arange(cdim, dtype=np.uint32), per_c)
    yj = np.concatenate([rng.choice(ycols, size=per_c, replace=False)
                         for _ in range(cdim)]).astype(np.uint32)
    Y = COOTensor((cdim, ycols), [yc, yj], (rng.random(len(yc)) + 0.1).astype(np.float32))
    return X, Y
