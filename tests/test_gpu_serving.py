"""GPU tests of the serving loop (sptc_server_*): every request's result is
the same served alone, served in a mix with other requests (batched and
overlapping a large request on another stream), and computed by the plain
single-call path."""
import time

import numpy as np
import pytest

from gsparc_amd import sptc
from helpers import random_coo

pytestmark = pytest.mark.gpu


def _tensors():
    rng = np.random.default_rng(77)
    small = []
    for k in range(7):
        dims = (int(rng.integers(20, 300)), int(rng.integers(20, 300)), int(rng.integers(5, 40)))
        small.append((random_coo(rng, dims, int(rng.integers(500, 20_000)),
                                 positive=bool(k % 2)), [2], [2]))
    narrow = random_coo(rng, (12, 10, 30), 2_000)  # dense-accumulator instance
    small.append((narrow, [2], [2]))
    large = random_coo(rng, (400, 300, 20), 200_000)  # ~2e9 products: served count-only
    return small, large


def _same(a, b):
    for x, y in zip(a.to_host(), b.to_host()):
        np.testing.assert_array_equal(x, y)


def test_results_same_alone_and_in_the_mix(gpu):
    small, large = _tensors()
    srv = sptc.Server(budget_bytes=24 << 30, workers=3, large_products=50_000_000,
                      large_budget_bytes=6 << 30)
    ids = [srv.add_tensor(t) for t, _, _ in small]
    big = srv.add_tensor(large)
    reqs = [(i, cx, cy) for i, (_, cx, cy) in zip(ids, small)]

    alone = []
    for tid, cx, cy in reqs:
        r, info = srv.wait(srv.submit(tid, tid, cx, cy))
        assert info["status"] == 0 and info["count_only"] == 0
        alone.append(r)
    rl, linfo = srv.wait(srv.submit(big, big, [2], [2]))
    assert linfo["count_only"] == 1
    big_alone = (rl.nnz, rl.info.value_sum, rl.ir_nnz)

    # each small result equals the plain single call
    for (t, cx, cy), r in zip(small, alone):
        p = sptc.SpTC(t, t, cx, cy)
        assert r.ir_nnz == p.ir_nnz and r.info.accumulator == p.info.accumulator
        _same(r, p)
    p = sptc.SpTC(large, large, [2], [2], count_only=True)
    assert p.nnz == big_alone[0] and p.ir_nnz == big_alone[2]

    # the mix: the large request first, then every small request three times
    order = [("big",)] + [reqs[k % len(reqs)] for k in range(3 * len(reqs))]
    sub = []
    for q in order:
        if q[0] == "big":
            sub.append(("big", srv.submit(big, big, [2], [2])))
        else:
            sub.append((q, srv.submit(q[0], q[0], q[1], q[2])))
    sizes = []
    for q, rid in sub:
        r, info = srv.wait(rid)
        sizes.append(info["batch_size"])
        if q == "big":
            assert (r.nnz, r.ir_nnz) == (big_alone[0], big_alone[2])
            assert abs(r.info.value_sum - big_alone[1]) <= 1e-12 * abs(big_alone[1])
        else:
            k = reqs.index(q)
            _same(r, alone[k])
            # value_sum is an FP64 checksum summed in no fixed order
            assert abs(r.info.value_sum - alone[k].info.value_sum) <= \
                1e-12 * abs(alone[k].info.value_sum)
        del r
    st = srv.stats()
    assert st["requests_done"] == len(reqs) + 1 + len(order)
    assert st["large_requests"] == 2
    assert st["peak_in_use_bytes"] <= st["budget_bytes"]
    assert max(sizes) > 1, sizes  # some requests were batched together
    srv.close()


def test_large_requests_run_one_at_a_time(gpu):
    """Large requests are HBM-bound: three submitted together, with budget
    and workers for all three, still run one after another (each one's
    launch starts after the previous one's result is ready)."""
    _, large = _tensors()
    srv = sptc.Server(budget_bytes=24 << 30, workers=3, large_products=50_000_000,
                      large_budget_bytes=6 << 30)
    big = srv.add_tensor(large)
    srv.wait(srv.submit(big, big, [2], [2]), want_result=False)  # operands prepared
    subs = []
    for _ in range(3):
        t0 = time.perf_counter() * 1e3
        subs.append((srv.submit(big, big, [2], [2]), t0))
    spans = []
    for rid, t0 in subs:
        _, info = srv.wait(rid, want_result=False)
        assert info["status"] == 0 and info["count_only"] == 1
        spans.append((t0 + info["queue_ms"], t0 + info["latency_ms"]))
    spans.sort()
    for (_, end0), (start1, _) in zip(spans, spans[1:]):
        assert start1 >= end0 - 2.0, spans  # ms; submit times are taken just before the call


def test_server_errors(gpu):
    rng = np.random.default_rng(5)
    t = random_coo(rng, (10, 10, 10), 100)
    srv = sptc.Server(budget_bytes=1 << 30)
    tid = srv.add_tensor(t)
    with pytest.raises(sptc.SpTCError):
        srv.submit(tid, tid + 5, [2], [2])
    with pytest.raises(sptc.SpTCError):
        srv.submit(tid, tid, [7], [2])
    r, info = srv.wait(srv.submit(tid, tid, [2], [2]))
    p = sptc.SpTC(t, t, [2], [2])
    _same(r, p)
    srv.close()
