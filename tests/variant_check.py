"""Runs the sort/reduce edge-case parity checks once in this process, for a
kernel variant chosen by environment variables set before the library
loads (test_gpu_variants.py starts one process per variant)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import test_gpu_parity as t  # noqa: E402
from gsparc_amd import sptc  # noqa: E402


def main():
    sptc.lib()
    t.test_row_segments_big_and_small_rows(None)
    for ycols in (1, 3, 40):
        t.test_long_runs_and_row_starts_across_wave_tiles(None, ycols)
    for n, bits, rb in ((100_000, 43, 8), (1_000_003, 48, 8), (700_000, 36, 9),
                        (3_000_000, 36, 0)):
        t.test_sort_pairs_stable_bit_exact(None, n, bits, rb)
    print("variant ok:", {k: v for k, v in os.environ.items() if k.startswith("SPTC_")},
          flush=True)


if __name__ == "__main__":
    main()
