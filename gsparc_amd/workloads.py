"""Synthetic SpTC workloads (seeded).

  plain_small   order-3 2000x2000x500, 2e5 nnz, Zipf(1.1) modes 0-1, uniform
                mode 2, uniform(0,1] values; Z = X x_{2,2} X.
  nips          order-4 2482x2862x14036x17, 3.1e6 nnz, Zipf(1.2) modes 0-2,
                uniform mode 3, lognormal(0,1) values; Z = X x_{{2,3},{2,3}} X.
  nell2         order-3 12092x9184x28818, 7.69e7 nnz, Zipf(1.1) every mode,
                uniform(0,1] values; Z = X x_{2,2} X under a 64 GB cap.
  batch         256 independent self-contractions, per-instance seed, each
                capped at 2^24 intermediate products (batch_specs).
  serving       Poisson request stream; 90 % batch-distribution requests,
                10 % large order-3 requests.

Seeds are 0x5EED0000 + workload number; instance i of a multi-instance
workload uses splitmix64(seed + i).
"""
import dataclasses
import functools
import math
from dataclasses import dataclass

import numpy as np

from .coo import (UNIFORM, ZIPF, VAL_LOGNORMAL, VAL_UNIFORM01, generate)

MASK64 = 2**64 - 1


def splitmix64(z):
    """Pure-Python SplitMix64 step (same as sptc_gen_splitmix64)."""
    z = (z + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


class _Stream:
    def __init__(self, seed):
        self.state = seed & MASK64
        self.k = 0

    def u64(self):
        self.k += 1
        return splitmix64((self.state + self.k * 0x632BE59BD9B4E019) & MASK64)

    def unit(self):
        return (self.u64() >> 11) * (1.0 / 9007199254740992.0)

    def log_uniform(self, lo, hi):
        return math.exp(math.log(lo) + self.unit() * (math.log(hi) - math.log(lo)))

    def choice(self, n):
        return (self.u64() * n) >> 64


@dataclass
class Spec:
    name: str
    seed: int
    dims: tuple
    dist: tuple
    zipf_s: tuple
    nnz: int
    value_dist: int
    cx: tuple
    cy: tuple

    def make(self):
        t = generate(self.seed, self.dims, self.dist, self.zipf_s, self.nnz,
                     self.value_dist, name=self.name)
        return t


SEED0 = 0x5EED0000

PLAIN_SMALL = Spec("plain_small_2000x2000x500_2e5", SEED0 + 0, (2000, 2000, 500),
                   (ZIPF, ZIPF, UNIFORM), (1.1, 1.1, 1.0), 200_000,
                   VAL_UNIFORM01, (2,), (2,))
NIPS = Spec("nips_2482x2862x14036x17_3.1e6", SEED0 + 1, (2482, 2862, 14036, 17),
            (ZIPF, ZIPF, ZIPF, UNIFORM), (1.2, 1.2, 1.2, 1.0), 3_100_000,
            VAL_LOGNORMAL, (2, 3), (2, 3))
NELL2 = Spec("nell2_12092x9184x28818_7.69e7", SEED0 + 2, (12092, 9184, 28818),
             (ZIPF, ZIPF, ZIPF), (1.1, 1.1, 1.1), 76_900_000,
             VAL_UNIFORM01, (2,), (2,))
NELL2_CAP_BYTES = 64 << 30


def batch_instance(i, seed=SEED0 + 3, nnz_cap=2_000_000):
    """Instance i of the 256-contraction batch: order 3-4, dims log-uniform in
    [100, 20000], nnz log-uniform in [1e3, 2e6] (capped at a quarter of the
    coordinate space), 1-2 contracted modes, Zipf(1.1) on every mode, values
    uniform(0,1]."""
    st = _Stream(splitmix64((seed + i) & MASK64))
    order = 3 + st.choice(2)
    dims = tuple(int(round(st.log_uniform(100, 20000))) for _ in range(order))
    space = 1
    for d in dims:
        space *= d
    nnz = int(round(st.log_uniform(1e3, nnz_cap)))
    nnz = min(nnz, space // 4)
    nc = 1 + st.choice(2)
    modes = list(range(order))
    cmodes = []
    for _ in range(nc):
        j = st.choice(len(modes))
        cmodes.append(modes.pop(j))
    cmodes.sort()
    return Spec(f"batch{i}", splitmix64(seed + 1000003 * (i + 1)), dims,
                (ZIPF,) * order, (1.1,) * order, nnz, VAL_UNIFORM01,
                tuple(cmodes), tuple(cmodes))


BATCH_IR_CAP = 1 << 24


def self_products(t, cmodes):
    """Intermediate products of T x_{c,c} T: sum over contract keys of the
    squared key count."""
    ck = np.zeros(t.nnz, np.int64)
    for m in cmodes:
        ck = ck * t.dims[m] + t.idx[m].astype(np.int64)
    _, cnt = np.unique(ck, return_counts=True)
    return int((cnt.astype(np.float64) ** 2).sum())


@functools.lru_cache(maxsize=4)
def batch_specs(n=256, ir_cap=BATCH_IR_CAP):
    """The first n batch instances with each one's nnz scaled down until its
    self-contraction has at most ir_cap intermediate products.  Drawn as
    stated, the distribution's products are heavy-tailed (256 instances:
    8.4e11 products in all, one instance alone 1.2e11, a quarter above
    2.7e8), so a few instances would be the whole workload; 104 of the 256
    are scaled."""
    out = []
    for i in range(n):
        s = batch_instance(i)
        for _ in range(8):
            ir = self_products(s.make(), s.cx)
            if ir <= ir_cap:
                break
            s = dataclasses.replace(s, nnz=max(1000, int(s.nnz * math.sqrt(ir_cap / ir) * 0.9)))
        out.append(s)
    return tuple(out)


def serving_large(i, seed=SEED0 + 4):
    """The 10 % large serving request: order 3, dims ~1e4, 1e7 nnz, one
    contracted mode.  Indices are uniform (pinned choice: a Zipf(1.1)
    contracted mode of extent 1e4 would give ~2e12 products per request)."""
    st = _Stream(splitmix64((seed + 7919 * (i + 1)) & MASK64))
    dims = tuple(int(round(st.log_uniform(8000, 12000))) for _ in range(3))
    c = st.choice(3)
    return Spec(f"large{i}", splitmix64(seed + 104729 * (i + 1)), dims,
                (UNIFORM,) * 3, (1.0,) * 3, 10_000_000, VAL_UNIFORM01, (c,), (c,))


CONFIGS = {
    "plain_small": PLAIN_SMALL,
    "nips": NIPS,
    "nell2": NELL2,
}
