"""GPU parity of the batched path (sptc_batch_*): every instance of a batch
must equal the golden fixtures and, bit for bit, the plain single-call path
on the same operands -- indices, values, product counts and accumulator
choice -- across eager replays and hipGraph replays."""
import numpy as np
import pytest

from gsparc_amd import sptc
from gsparc_amd.coo import COOTensor
from gsparc_amd.workloads import batch_specs
from helpers import REL_TOL, assert_same_coo, dense_friendly, load_fixtures, random_coo

pytestmark = pytest.mark.gpu
FIXTURES = load_fixtures()


def _same(a, b):
    for x, y in zip(a.to_host(), b.to_host()):
        np.testing.assert_array_equal(x, y)


def _mixed_items(seed):
    """Uneven instances: orders 2-4, 0-2 contracted modes, small and wide
    column spaces (dense and ESC), duplicates, an empty operand, X != Y."""
    rng = np.random.default_rng(seed)
    items = []
    for k in range(14):
        xo = int(rng.integers(2, 5))
        nc = int(rng.integers(0 if k == 3 else 1, min(xo, 3)))
        yo = int(rng.integers(max(nc, 2), 5))
        cdims = [int(x) for x in rng.integers(2, 40, size=nc)]
        small = k % 3 == 0  # narrow free extents favour the dense accumulator
        hi = 12 if small else 400
        xdims = [int(x) for x in rng.integers(2, hi, size=xo)]
        ydims = [int(x) for x in rng.integers(2, hi, size=yo)]
        cx = [int(x) for x in rng.permutation(xo)[:nc]]
        cy = [int(x) for x in rng.permutation(yo)[:nc]]
        for a, b, d in zip(cx, cy, cdims):
            xdims[a] = d
            ydims[b] = d
        X = random_coo(rng, xdims, int(rng.integers(1, 6000)), positive=bool(k % 2))
        Y = random_coo(rng, ydims, int(rng.integers(1, 6000)), positive=bool(k % 2))
        if k == 5:  # duplicate Y coordinates (ESC only)
            Y = COOTensor(Y.dims, [np.concatenate([a, a[:50]]) for a in Y.idx],
                          np.concatenate([Y.val, Y.val[:50] * 0.5]))
        if k == 7:  # empty X
            X = COOTensor(X.dims, [a[:0] for a in X.idx], X.val[:0])
        if k == 9:
            Y = X if xo == yo and cx == cy else Y
        items.append((X, Y, cx, cy))
    # few owners per row over long Y segments: the automatic rule's dense shape
    X, Y = dense_friendly(rng)
    items.append((X, Y, [1], [0]))
    return items


def test_batch_matches_golden_fixtures(gpu):
    b = sptc.Batch([(c["X"], c["Y"], c["cx"], c["cy"]) for c in FIXTURES])
    for i, case in enumerate(FIXTURES):
        r = b.result(i)
        idx, v = r.decode()
        assert_same_coo([a.astype(np.int64) for a in idx], v, case["zcoords"], case["zval"],
                        rtol=REL_TOL)


@pytest.mark.parametrize("dense", [0, -1])
@pytest.mark.parametrize("seed", [1, 2])
def test_batch_equals_plain_bitwise(gpu, seed, dense):
    items = _mixed_items(seed)
    b = sptc.Batch(items, dense_accumulator=dense)
    info = b.info
    assert info["instances"] == len(items)
    tot = 0
    for i, (X, Y, cx, cy) in enumerate(items):
        p = sptc.SpTC(X, Y, cx, cy, dense_accumulator=dense)
        r = b.result(i)
        assert r.ir_nnz == p.ir_nnz and r.nnz == p.nnz
        assert r.info.nrows == p.info.nrows
        assert r.info.accumulator == p.info.accumulator
        _same(r, p)
        np.testing.assert_array_equal(r.row_keys(), p.row_keys())
        tot += p.nnz
    assert info["out_nnz"] == tot
    if dense == 0:
        assert info["dense_instances"] > 0


def test_batch_chunked_pieces_equal_plain(gpu):
    """A budget that splits the ESC products into several chunks, each
    spanning pieces of several instances."""
    rng = np.random.default_rng(11)
    items = []
    for k in range(6):
        X = random_coo(rng, (50 + 7 * k, 40, 30), 20_000)
        items.append((X, X, [2], [2]))
    plain = [sptc.SpTC(X, Y, cx, cy, dense_accumulator=-1) for X, Y, cx, cy in items]
    ir = sum(p.ir_nnz for p in plain)
    out = sum(p.nnz for p in plain)
    b = sptc.Batch(items, dense_accumulator=-1,
                   mem_budget_bytes=out * 16 + ir * 10 + (256 << 20))
    assert b.info["esc_chunks"] > 1
    for i, p in enumerate(plain):
        r = b.result(i)
        assert r.info.nchunks >= 1
        _same(r, p)
    # row ranges through the instance views
    r = b.result(2)
    rk = r.row_keys()
    for j in (0, len(rk) // 2, len(rk) - 1):
        beg, end = r.row_range(j)
        pb, pe = plain[2].row_range(j)
        assert (beg, end) == (pb, pe)


def _device_items(torch, items):
    class Dev:
        pass

    out = []
    for X, Y, cx, cy in items:
        dx = Dev()
        dx.dims, dx.order, dx.nnz = X.dims, X.order, X.nnz
        dx.idx = [torch.from_numpy(a.view(np.int32)).cuda() for a in X.idx]
        dx.val = torch.from_numpy(X.val).cuda()
        if Y is X:
            dy = dx
        else:
            dy = Dev()
            dy.dims, dy.order, dy.nnz = Y.dims, Y.order, Y.nnz
            dy.idx = [torch.from_numpy(a.view(np.int32)).cuda() for a in Y.idx]
            dy.val = torch.from_numpy(Y.val).cuda()
        out.append((dx, dy, cx, cy))
    return out


def test_batch_replay_and_graph(gpu):
    """Eager and graph replays, with other GPU work between them, reproduce
    the planned results; a replay re-reads operand values."""
    import torch
    items = [it for it in _mixed_items(3) if it[0].nnz and it[1].nnz]
    ditems = _device_items(torch, items)
    b = sptc.Batch(ditems, on_device=True)
    base = [b.result(i).to_host() for i in range(len(items))]
    X0, Y0 = items[0][0], items[0][1]
    for use_graph in (False, True, True, False, True):
        junk = sptc.SpTC(X0, Y0, items[0][2], items[0][3])  # eager work between replays
        t = torch.randn(1 << 20, device="cuda")
        (t * 2).sum().item()
        del junk
        b.run(use_graph=use_graph)
        b.sync()
        for i in range(len(items)):
            for x, y in zip(b.result(i).to_host(), base[i]):
                np.testing.assert_array_equal(x, y)
    assert b.info["graph_nodes"] > 0
    # doubling every Y value doubles every output exactly
    for _, dy, _, _ in ditems:
        dy.val.mul_(2.0)
    torch.cuda.synchronize()
    b.run(use_graph=True)
    b.sync()
    for i in range(len(items)):
        xf, yf, v = b.result(i).to_host()
        np.testing.assert_array_equal(xf, base[i][0])
        np.testing.assert_array_equal(yf, base[i][1])
        np.testing.assert_array_equal(v, base[i][2] * np.float32(2.0))


def test_batch_stale_plan_is_reported(gpu):
    """Changing an index array after planning is detected before any
    plan-sized work runs; the results stay as they were."""
    import torch
    rng = np.random.default_rng(21)
    items = [(random_coo(rng, (30, 20, 10), 800), random_coo(rng, (10, 25), 200), [2], [0])
             for _ in range(3)]
    ditems = _device_items(torch, items)
    b = sptc.Batch(ditems, on_device=True)
    before = b.result(1).to_host()
    ditems[1][0].idx[0].copy_(torch.flip(ditems[1][0].idx[0], [0]))
    torch.cuda.synchronize()
    b.run(use_graph=False)
    with pytest.raises(sptc.SpTCError) as e:
        b.sync()
    assert e.value.status == sptc.ERR_INTERNAL
    for x, y in zip(b.result(1).to_host(), before):
        np.testing.assert_array_equal(x, y)


def test_batch_distribution_sample(gpu):
    """The first 24 instances of the batched workload (order 3-4, 1-2
    contracted modes, Zipf indices) equal their plain calls."""
    specs = batch_specs(24)
    items = []
    for s in specs:
        t = s.make()
        items.append((t, t, list(s.cx), list(s.cy)))
    b = sptc.Batch(items)
    for i, (X, Y, cx, cy) in enumerate(items):
        p = sptc.SpTC(X, Y, cx, cy)
        r = b.result(i)
        assert r.info.accumulator == p.info.accumulator
        _same(r, p)


def test_batch_errors(gpu):
    rng = np.random.default_rng(4)
    X = random_coo(rng, (10, 10, 10), 50)
    with pytest.raises(sptc.SpTCError) as e:
        sptc.Batch([(X, X, [2], [2]), (X, X, [5], [2])])
    assert e.value.status == sptc.ERR_BAD_ARGS
    b = sptc.Batch([(X, X, [2], [2])])
    with pytest.raises(sptc.SpTCError):
        b.result(1)
