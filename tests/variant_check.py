"""Runs the sort/reduce edge-case parity checks once in this process, for a
kernel variant chosen by environment variables set before the library
loads (test_gpu_variants.py starts one process per variant)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import test_gpu_batch as tb  # noqa: E402
import test_gpu_parity as t  # noqa: E402
from gsparc_amd import sptc  # noqa: E402
from helpers import random_coo  # noqa: E402


def dense_checks():
    """Dense accumulators (single and batched) against the oracle and the
    golden fixtures, to the value tolerance (the FP64 LDS-atomic form is not
    bit-identical to ESC)."""
    for case in t.SMALL_COL:
        t.test_dense_matches_golden(None, case)
    for seed in range(2):
        rng = np.random.default_rng(500 + seed)
        X = random_coo(rng, (400, 300, 50), 80_000, positive=bool(seed % 2))
        Y = random_coo(rng, (50, 60, 90), 60_000, positive=bool(seed % 2))
        d = sptc.SpTC(X, Y, [2], [0], dense_accumulator=1)
        assert d.info.accumulator == 1
        t._compare_with_oracle(X, Y, [2], [0], d)
    tb.test_batch_matches_golden_fixtures(None)


def digests():
    """One line per big-row case: a digest of the output's indices and value
    bits, for comparing two variants bit for bit."""
    import hashlib
    from gsparc_amd.coo import UNIFORM, ZIPF, VAL_UNIFORM01, generate
    cases = [t.bucket_case(n) + ([1], [0]) for n in t.BUCKET_CASES]
    X = generate(0x5EED00AA, (300, 200, 48), (ZIPF, ZIPF, UNIFORM), (1.3, 1.3, 1.0), 40_000,
                 VAL_UNIFORM01)
    cases.append((X, X, [2], [2]))
    for i, (X, Y, cx, cy) in enumerate(cases):
        res = sptc.SpTC(X, Y, cx, cy, dense_accumulator=-1)
        h = hashlib.sha256()
        for a in res.to_host():
            h.update(np.ascontiguousarray(a).tobytes())
        print(f"digest {i} {res.nnz} {h.hexdigest()}", flush=True)


def main():
    sptc.lib()
    if "--digest" in sys.argv:
        digests()
    elif os.environ.get("SPTC_DENSE_ORDER") == "atomic":
        dense_checks()
    else:
        t.test_row_segments_big_and_small_rows(None)
        for name in t.BUCKET_CASES:
            t.test_big_rows_wide_columns_against_oracle(None, name)
        for ycols in (1, 3, 40):
            t.test_long_runs_and_row_starts_across_wave_tiles(None, ycols)
        for n, bits, rb in ((100_000, 43, 8), (1_000_003, 48, 8), (700_000, 36, 9),
                            (3_000_000, 36, 0)):
            t.test_sort_pairs_stable_bit_exact(None, n, bits, rb)
    print("variant ok:", {k: v for k, v in os.environ.items() if k.startswith("SPTC_")},
          flush=True)


if __name__ == "__main__":
    main()
