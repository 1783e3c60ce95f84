"""Writes sptc_fixtures.json: small SpTC cases with their exact outputs.

Inputs come from the seeded generator (gsparc_amd/csrc/gen.cpp); expected
outputs come from a dense numpy.tensordot of the same tensors in FP64 --
independent of the oracle and of the GPU path -- listed as (output
coordinates, value) for every nonzero, output modes ordered X free modes
ascending, then Y free modes ascending.  Run from the repo root:
    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from gsparc_amd.coo import UNIFORM, ZIPF, generate  # noqa: E402

CASES = [
    # name, (X dims, X nnz, seed), (Y dims, Y nnz, seed) or None for X, cx, cy
    ("self_o3_c2", ((6, 7, 5), 60, 11), None, [2], [2]),
    ("xy_o3_c21_c02", ((5, 6, 7), 70, 12), ((7, 4, 6), 50, 13), [2, 1], [0, 2]),
    ("self_o4_c23", ((4, 5, 3, 6), 90, 14), None, [2, 3], [2, 3]),
    ("spgemm_o2", ((8, 9), 30, 15), ((9, 7), 25, 16), [1], [0]),
    ("scalar_full", ((4, 5), 12, 17), ((4, 5), 9, 18), [0, 1], [0, 1]),
    ("outer_nc0", ((3, 4), 5, 19), ((2, 5), 4, 20), [], []),
    ("dim1_mode", ((1, 6, 5), 20, 21), None, [2], [2]),
    ("pow2_dims", ((8, 16, 4), 100, 22), None, [0], [0]),
    ("mixed_order_o4_o3", ((3, 4, 5, 2), 50, 23), ((5, 3, 6), 40, 24), [2, 0], [0, 1]),
    ("zipf_heavy", ((9, 10, 6), 150, 25), None, [1], [1]),
]


def make(dims, nnz, seed):
    dist = [ZIPF] * len(dims)
    return generate(seed, dims, dist, [1.1] * len(dims), nnz)


def dense(t):
    D = np.zeros(t.dims, dtype=np.float64)
    D[tuple(a.astype(np.int64) for a in t.idx)] = t.val.astype(np.float64)
    return D


def main():
    out = []
    for name, xs, ys, cx, cy in CASES:
        X = make(*xs)
        Y = make(*ys) if ys else X
        Z = np.tensordot(dense(X), dense(Y), axes=(cx, cy))
        nzi = np.nonzero(Z) if Z.ndim else ((),)
        if Z.ndim == 0:
            coords, vals = [], [float(Z)] if Z != 0 else []
        else:
            coords = [a.tolist() for a in nzi]
            vals = Z[nzi].tolist()
        out.append({
            "name": name, "cx": cx, "cy": cy,
            "X": {"dims": list(X.dims), "idx": [a.tolist() for a in X.idx],
                  "val": [float(v) for v in X.val]},
            "Y": None if ys is None else {"dims": list(Y.dims), "idx": [a.tolist() for a in Y.idx],
                                          "val": [float(v) for v in Y.val]},
            "Z": {"order": int(Z.ndim), "dims": list(Z.shape), "coords": coords,
                  "val": vals},
        })
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "sptc_fixtures.json"), "w") as f:
        json.dump({"source": "numpy.tensordot over dense FP64 copies", "cases": out}, f)


if __name__ == "__main__":
    main()
