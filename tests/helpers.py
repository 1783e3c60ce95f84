"""Shared helpers for the SpTC tests: fixtures, key decoding, comparisons."""
import json
import os

import numpy as np

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
