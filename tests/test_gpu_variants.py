"""Kernel variants selected by environment variables (read once per
process): each runs the sort and reduce edge-case parity checks against the
oracle in a process of its own, one after another.

  SPTC_SEG_WIDE=1      1024-thread x 8-key big-row scatter tiles
  SPTC_SEG_IPT=12/24   keys per thread of the big-row scatter tiles
  SPTC_SEG_RB=8        8-bit digits for the big-row column passes
  SPTC_SORT=onesweep   operand sorts with a decoupled look-back scan
  SPTC_SORT_TILE=256x16/512x16  operand-sort tile (threads x keys, 8-bit digits)
  SPTC_SORT_RB9=0      operand sorts keep 8-bit digits where 9 would save a pass
  SPTC_PROD_SORT=full  one sort over all product-key bits (64-bit keys)
  SPTC_XCD_REMAP=0     blockIdx-order tiles instead of XCD-contiguous ones
  SPTC_DENSE_ORDER=atomic  dense rows summed by FP64 LDS atomics (checked
                       against the oracle and the golden fixtures instead)
  SPTC_BIG_TAIL=buckets  big rows of 11 to 20 column bits: one scatter pass
                       on the top 10 bits, then an LDS reduce per bucket
                       (also compared bit for bit with the default)
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = [{"SPTC_SEG_WIDE": "1"}, {"SPTC_SEG_IPT": "12"}, {"SPTC_SEG_IPT": "24"},
            {"SPTC_SEG_RB": "8"}, {"SPTC_SORT": "onesweep"}, {"SPTC_SORT_TILE": "256x16"},
            {"SPTC_SORT_TILE": "512x16"}, {"SPTC_SORT_RB9": "0"}, {"SPTC_PROD_SORT": "full"},
            {"SPTC_XCD_REMAP": "0"}, {"SPTC_DENSE_ORDER": "atomic"},
            {"SPTC_BIG_TAIL": "buckets"}]


@pytest.mark.parametrize("env", VARIANTS, ids=lambda e: ",".join(f"{k}={v}" for k, v in e.items()))
def test_variant_against_oracle(gpu, env):
    print(f"[variant] {env}", flush=True)
    p = subprocess.run([sys.executable, "-u", os.path.join(HERE, "variant_check.py")],
                       env=dict(os.environ, **env), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=150)
    assert p.returncode == 0, p.stdout[-3000:]
    assert "variant ok" in p.stdout


def test_big_row_buckets_equal_the_full_column_sort(gpu):
    """The high-digit bucket reduce (SPTC_BIG_TAIL=buckets) and the default
    sort on all column bits sum every output's products in the same order
    with the same steps: identical outputs, bit for bit, on big-row cases of
    11 to 20 column bits."""
    outs = []
    for env in ({"SPTC_BIG_TAIL": "buckets"}, {"SPTC_BIG_TAIL": "columns"}):
        print(f"[variant] digests {env}", flush=True)
        p = subprocess.run([sys.executable, "-u", os.path.join(HERE, "variant_check.py"),
                            "--digest"], env=dict(os.environ, **env), stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=150)
        assert p.returncode == 0, p.stdout[-3000:]
        outs.append([ln for ln in p.stdout.splitlines() if ln.startswith("digest")])
    assert len(outs[0]) == 4
    assert outs[0] == outs[1]
