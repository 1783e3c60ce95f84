"""The oracle against the reference's own CPU SpTC: Sparta `ttt`, built
from the reference's sources into oracle/_ref/ by oracle/Makefile.  Output
coordinates must be the same set; values agree to Sparta's printed
precision (it sums float-rounded products in double, tmul.c:1206, and
prints 6 decimals, dump.c:49)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from gsparc_amd.coo import COOTensor, write_tns
from helpers import decode_x, decode_y, random_coo

pytestmark = pytest.mark.skipif(not os.path.exists(orc.REF_TTT),
                                reason="reference CPU SpTC not built (oracle/_ref/ttt)")


def _sparta(X, Y, cx, cy, tmp_path):
    px, py, pz = (str(tmp_path / n) for n in ("x.tns", "y.tns", "z.txt"))
    write_tns(px, X)
    write_tns(py, Y)
    cmd = [orc.REF_TTT, "-X", px, "-Y", py, "-m", str(len(cx)), "-x", *map(str, cx),
           "-y", *map(str, cy), "-t", "1", "-Z", pz]
    p = subprocess.run(cmd, env=dict(os.environ, EXPERIMENT_MODES="3"), stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    with open(pz) as f:
        lines = f.read().split("\n")
    rows = [ln.split() for ln in lines[2:] if ln.strip()]
    coords = np.array([[int(t) for t in r[:-1]] for r in rows], dtype=np.int64)
    vals = np.array([float(r[-1]) for r in rows], dtype=np.float64)
    return coords, vals


def _oracle(X, Y, cx, cy):
    r = orc.pipeline(X, Y, cx, cy)
    sx, wx, fx, _, ncx = r["xl"]
    sy, wy, fy, _, _ = r["yl"]
    cols = decode_x(r["xf"], sx, wx, fx, ncx) + decode_y(r["yf"], sy, wy, fy)
    return np.stack([c.astype(np.int64) for c in cols], axis=1), r["v64"]


@pytest.mark.parametrize("case", [
    ((30, 25, 12), 3000, (12, 20, 15), 1500, [2], [0]),
    ((20, 15, 8, 6), 4000, (8, 6, 18), 700, [2, 3], [0, 1]),
    ((40, 30, 10), 5000, None, None, [2], [2]),
])
def test_oracle_matches_reference_cpu_sptc(case, tmp_path):
    xd, xn, yd, yn, cx, cy = case
    rng = np.random.default_rng(sum(xd) + xn)
    X = random_coo(rng, xd, xn)
    Y = X if yd is None else random_coo(rng, yd, yn)
    # the .tns reader takes dims from the observed maximum index
    X = COOTensor(tuple(int(a.max()) + 1 for a in X.idx), X.idx, X.val)
    Y = X if yd is None else COOTensor(tuple(int(a.max()) + 1 for a in Y.idx), Y.idx, Y.val)
    sc, sv = _sparta(X, Y, cx, cy, tmp_path)
    oc, ov = _oracle(X, Y, cx, cy)
    assert len(sc) == len(oc) > 0
    so = np.lexsort(sc.T[::-1])
    oo = np.lexsort(oc.T[::-1])
    np.testing.assert_array_equal(sc[so], oc[oo])
    np.testing.assert_allclose(sv[so], ov[oo], rtol=1e-5, atol=2e-6)
