"""SpTC benchmark on one MI355X.

A step is one full SpTC -- linearize, sort, index match, partition and ESC
accumulation -- of the NIPS-shaped workload (order-4 2482x2862x14036x17,
3.1e6 nnz, Zipf(1.2) modes 0-2, lognormal values) contracted with itself over
modes {2,3}, inputs resident in HBM when the timed region starts, output left
resident in HBM (freed after the step).

Prints one JSON line: output nnz/s plus the roofline of the dominant kernel
and the reference CPU SpTC (Sparta ttt) timed on a bounded sample.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="nips",
                   choices=["nips", "plain_small", "nell2", "batch", "batch_stated",
                            "serving"])
    p.add_argument("--requests", type=int, default=200)
    p.add_argument("--nell2-full", action="store_true",
                   help="nell2: the whole contraction count-only (about 9 minutes) "
                        "instead of a row sample")
    p.add_argument("--nell2-products", type=float, default=1e11,
                   help="intermediate products of the nell2 row sample")
    p.add_argument("--qps", type=float, default=10.0)
    p.add_argument("--batch-instances", type=int, default=256)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample-products", type=float, default=0,
                   help="intermediate products in the CPU-baseline sample (0: auto)")
    p.add_argument("--radix-bits", type=int, default=0, help="0: autotune, 8 or 11")
    p.add_argument("--accumulator", type=int, default=0,
                   help="dense_accumulator: 0 planner's choice, -1 ESC, 1 dense, 2 tiled dense")
    p.add_argument("--copy-probe-bytes", type=int, default=0,
                   help="run one device-to-device copy of this size first (counter calibration)")
    return p.parse_args()


def to_device(t):
    import torch

    class Dev:
        pass
    d = Dev()
    d.dims = t.dims
    d.order = t.order
    d.idx = [torch.from_numpy(a.view(np.int32)).cuda() for a in t.idx]
    d.val = torch.from_numpy(t.val).cuda()
    d.nnz = t.nnz
    return d


def measured_traffic(workload, kernel_class):
    """HBM bytes per launch of the dominant kernel class, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (calibrated
    against a copy of known size; see the summary), or None."""
    path = os.path.join(HERE, "profiles", f"{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"^bench-traffic: (\w+) ([0-9.e+]+)\s*$", line)
        if m and m.group(1) == kernel_class:
            return float(m.group(2))
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, X, target_products):
    """Reference CPU SpTC (Sparta ttt, EXPERIMENT_MODES=3) on a bounded
    sample: the nonzeros of every k-th X free row (rows in free-key order, k
    chosen so the sample holds about target_products intermediate
    products, so it spans the heavy head and the light tail of the row
    distribution) contracted against all of X."""
    import oracle as orc
    from gsparc_amd.coo import COOTensor, write_tns_fast
    if not os.path.exists(orc.REF_TTT):
        return None
    free = [m for m in range(X.order) if m not in spec.cx]
    fk = np.zeros(X.nnz, np.int64)
    for m in free:
        fk = fk * X.dims[m] + X.idx[m].astype(np.int64)
    ck = np.zeros(X.nnz, np.int64)
    for m in spec.cx:
        ck = ck * X.dims[m] + X.idx[m].astype(np.int64)
    _, cinv = np.unique(ck, return_inverse=True)
    per_nnz = np.bincount(cinv).astype(np.int64)[cinv]  # products of each nonzero
    rows, rinv = np.unique(fk, return_inverse=True)
    total = int(per_nnz.sum())
    stride = max(1, int(round(total / target_products)))
    sel = (rinv % stride) == 0
    Xs = COOTensor(X.dims, [a[sel] for a in X.idx], X.val[sel])
    prods = int(per_nnz[sel].sum())
    pick = rows[::stride]
    threads = os.cpu_count() or 1
    with tempfile.TemporaryDirectory() as d:
        xs_path = os.path.join(d, "xs.tns")
        x_path = os.path.join(d, "x.tns")
        write_tns_fast(xs_path, Xs)
        write_tns_fast(x_path, X)
        cmd = [orc.REF_TTT, "-X", xs_path, "-Y", x_path, "-m", str(len(spec.cx)), "-x",
               *[str(c) for c in spec.cx], "-y", *[str(c) for c in spec.cy], "-t", str(threads)]
        env = dict(os.environ, EXPERIMENT_MODES="3")
        p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=600)
    tot = None
    for line in p.stderr.splitlines():
        if line.startswith("[Total time]"):
            tot = float(line.split(":")[1].split()[0])
    nnz_z = re.findall(r"NNZ=(\d+)", p.stdout)  # the output tensor's status is last
    if tot is None or not nnz_z:
        return None
    return {"total_s": tot, "threads": threads, "sample_rows": int(len(pick)),
            "stride": stride, "rows": int(len(rows)),
            "sample_nnz": int(sel.sum()), "sample_products": prods,
            "sample_out_nnz": int(nnz_z[-1]), "cpu_model": cpu_model()}


def bench_batch(args):
    """256 independent self-contractions (batch_specs): one step is every
    instance once -- as 256 plain calls, and as one batch replayed through
    its captured hipGraph.  Output nnz/s of the batched form is the value."""
    import torch
    from gsparc_amd import sptc
    from gsparc_amd.workloads import BATCH_IR_CAP, batch_specs

    log("generating the batch instances")
    specs = batch_specs(args.batch_instances)
    tens = [s.make() for s in specs]
    dev = [to_device(t) for t in tens]
    log("plain calls")
    items = [(d, d, list(s.cx), list(s.cy)) for d, s in zip(dev, specs)]
    torch.cuda.synchronize()

    def plain_step():
        tot = 0
        for d, _, cx, cy in items:
            r = sptc.SpTC(d, d, cx, cy, on_device=True)
            tot += r.nnz
            del r
        torch.cuda.synchronize()
        return tot

    for _ in range(args.warmup):
        plain_out = plain_step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain_out = plain_step()
    plain_s = (time.perf_counter() - t0) / args.steps

    log("batched")
    t0 = time.perf_counter()
    b = sptc.Batch(items, on_device=True)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    info = b.info
    # kernel breakdown of one planning run (per-launch events; not timed)
    bt = sptc.Batch(items, on_device=True, kernel_timing=True)
    bt_stats = bt.kernel_stats()
    kern = {k: {"launches": n, "total_ms": ms, "achieved_gbs": (by / (ms * 1e-3) / 1e9)
                if ms else None}
            for k, (n, ms, by) in bt_stats.items() if n}
    del bt
    if info["out_nnz"] != plain_out:
        raise SystemExit(f"batched output nnz {info['out_nnz']} != plain {plain_out}")
    res = {}
    for mode, use_graph in (("eager", False), ("graph", True)):
        for _ in range(max(1, args.warmup)):
            b.run(use_graph=use_graph)
            b.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run(use_graph=use_graph)
        b.sync()
        torch.cuda.synchronize()
        res[mode] = (time.perf_counter() - t0) / args.steps
    ir = info["ir_nnz"]
    out = {
        "metric": "SpTC output nnz/s (batched)",
        "value": plain_out / res["graph"],
        "unit": "nnz/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["graph"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch{len(specs)}_cap{BATCH_IR_CAP}",
                   "instances": len(specs), "nnz_in": info["nnz_in"], "ir_nnz": ir,
                   "out_nnz": plain_out},
        "plain_ms_per_step": plain_s * 1e3,
        "batched_eager_ms_per_step": res["eager"] * 1e3,
        "batched_plan_ms": plan_s * 1e3,
        "speedup_vs_plain": plain_s / res["graph"],
        "speedup_eager_vs_plain": plain_s / res["eager"],
        "batch_info": info,
        # launches per run: the captured graph's nodes (one per kernel or copy)
        "graph_nodes": b.info["graph_nodes"],
        "plan_run_kernels": kern,
    }
    peak = sptc.device_info()["peak_hbm_gbs"]
    ks = {k: v for k, v in bt_stats.items() if v[0] and v[1]}
    if ks:
        dom = max(ks, key=lambda k: ks[k][1])
        n_l, ms_l, by_l = ks[dom]
        ach = by_l / (ms_l * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": peak,
                           "unit": "GB/s", "frac": ach / peak, "traffic": None,
                           "avg_launch_ms": ms_l / n_l, "alg_bytes_per_launch": by_l / n_l}
    if not args.no_cpu_baseline:
        cb = cpu_baseline_instances(specs, tens, budget_s=20.0)
        if cb:
            out["cpu_baseline"] = cb
    print(json.dumps(out))


def cpu_baseline_instances(specs, tens, budget_s=20.0):
    """The reference's CPU SpTC (Sparta ttt) on whole instances of a batch,
    in instance order, until about budget_s seconds of CPU time are spent;
    output nnz/s over the instances run."""
    import oracle as orc
    from gsparc_amd.coo import write_tns_fast
    if not os.path.exists(orc.REF_TTT):
        return None
    threads = os.cpu_count() or 1
    tot_s, out_nnz, prods, done = 0.0, 0, 0, 0
    with tempfile.TemporaryDirectory() as d:
        for s_, t in zip(specs, tens):
            path = os.path.join(d, "x.tns")
            write_tns_fast(path, t)
            cmd = [orc.REF_TTT, "-X", path, "-Y", path, "-m", str(len(s_.cx)), "-x",
                   *[str(c) for c in s_.cx], "-y", *[str(c) for c in s_.cy], "-t", str(threads)]
            p = subprocess.run(cmd, env=dict(os.environ, EXPERIMENT_MODES="3"),
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                               timeout=600)
            tt = [float(l.split(":")[1].split()[0]) for l in p.stderr.splitlines()
                  if l.startswith("[Total time]")]
            nz = re.findall(r"NNZ=(\d+)", p.stdout)
            if not tt or not nz:
                return None
            tot_s += tt[0]
            out_nnz += int(nz[-1])
            done += 1
            if tot_s >= budget_s:
                break
    return {"value": out_nnz / tot_s if tot_s else None, "unit": "nnz/s", "cores": threads,
            "kind": "reference", "cpu_model": cpu_model(),
            "sample": f"Sparta ttt (EXPERIMENT_MODES=3, [Total time]) on the first {done} "
                      f"instances, one after another: {out_nnz} output nnz in {tot_s:.3f} s"}


def bench_batch_stated(args):
    """The 256 batch instances as drawn (no nnz scaling): instances with at
    most 2^28 products run as one batch (graph replay), the larger ones as
    plain calls (count-only above 2^31 products, whose outputs would not fit
    the device); against 256 plain calls.  Total output nnz/s."""
    import torch
    from gsparc_amd import sptc
    from gsparc_amd.workloads import batch_instance, self_products

    log("generating the instances as drawn")
    specs = [batch_instance(i) for i in range(args.batch_instances)]
    tens = [s.make() for s in specs]
    irs = [self_products(t, s.cx) for s, t in zip(specs, tens)]
    dev = [to_device(t) for t in tens]
    small = [i for i, ir in enumerate(irs) if ir <= (1 << 28)]
    large = [i for i, ir in enumerate(irs) if ir > (1 << 28)]
    torch.cuda.synchronize()

    def plain(i):
        d, s_ = dev[i], specs[i]
        return sptc.SpTC(d, d, list(s_.cx), list(s_.cy), on_device=True,
                         count_only=irs[i] > (1 << 31))

    def plain_all(ids):
        tot = 0
        for i in ids:
            r = plain(i)
            tot += r.nnz
            del r
        torch.cuda.synchronize()
        return tot

    plain_all(small)
    log("plain calls (all instances)")
    t0 = time.perf_counter()
    out_small = plain_all(small)
    t_small_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    out_large = plain_all(large)
    t_large = time.perf_counter() - t0
    log("batched (instances up to 2^28 products)")
    b = sptc.Batch([(dev[i], dev[i], list(specs[i].cx), list(specs[i].cy)) for i in small],
                   on_device=True)
    if b.info["out_nnz"] != out_small:
        raise SystemExit(f"batched output nnz {b.info['out_nnz']} != plain {out_small}")
    for _ in range(max(1, args.warmup)):
        b.run(use_graph=True)
        b.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run(use_graph=True)
    b.sync()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    total_out = out_small + out_large
    t_served = t_batch + t_large
    out = {
        "metric": "SpTC output nnz/s (batched, instances as drawn)",
        "value": total_out / t_served, "unit": "nnz/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_served * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch{len(specs)}_as_drawn", "instances": len(specs),
                   "batched_instances": len(small), "plain_instances": len(large),
                   "count_only_instances": sum(irs[i] > (1 << 31) for i in large),
                   "ir_nnz": int(sum(irs)), "out_nnz": total_out},
        "batched_ms": t_batch * 1e3, "batched_plain_ms": t_small_plain * 1e3,
        "batched_speedup": t_small_plain / t_batch,
        "large_plain_ms": t_large * 1e3,
        "all_plain_ms": (t_small_plain + t_large) * 1e3,
        "speedup_vs_plain": (t_small_plain + t_large) / t_served,
    }
    print(json.dumps(out))


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _percentiles(lat):
    a = np.sort(np.asarray(lat, np.float64))
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean()), "max_ms": float(a.max())}


def bench_serving(args):
    """Poisson request stream over resident tensors: 90 % requests drawn from
    the batched workload's instances, 10 % large (order 3, dims ~1e4, 1e7 nnz,
    one contracted mode; count-only responses).  The server (continuous
    batching, HBM-footprint admission under 96 GiB, resident prepared
    operands) against serial plain calls on the same arrival times, plus a
    burst (all requests at once) for throughput."""
    import threading
    import torch
    from gsparc_amd import sptc
    from gsparc_amd.workloads import batch_specs, serving_large

    log("generating tensors")
    specs = batch_specs(256)
    small = [s.make() for s in specs]
    larges = [serving_large(i) for i in range(4)]
    big = [s.make() for s in larges]
    log("server: registering and warming")
    rng = np.random.default_rng(0x5EED0004)
    n = args.requests
    kinds = rng.random(n) < 0.1
    picks = [int(rng.integers(len(big))) if k else int(rng.integers(len(small))) for k in kinds]
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, size=n))
    reqs = []
    for k, i in zip(kinds, picks):
        if k:
            reqs.append(("large", i, list(larges[i].cx)))
        else:
            reqs.append(("small", i, list(specs[i].cx)))

    # ---- server ------------------------------------------------------------
    budget = 96 << 30
    srv = sptc.Server(budget_bytes=budget, workers=3, large_products=1 << 28,
                      large_budget_bytes=budget // 3)
    sid = [srv.add_tensor(t) for t in small]
    bid = [srv.add_tensor(t) for t in big]

    def tid(q):
        return bid[q[1]] if q[0] == "large" else sid[q[1]]

    # warm the resident operand cache and the product counts, one request each
    seen = set()
    for q in reqs:
        key = (q[0], q[1])
        if key in seen:
            continue
        seen.add(key)
        r, _ = srv.wait(srv.submit(tid(q), tid(q), q[2], q[2]), want_result=False)
    torch.cuda.synchronize()

    def run_server(times):
        ids = [None] * n
        t0 = time.perf_counter()

        def submitter():
            for k, q in enumerate(reqs):
                dt = times[k] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids[k] = srv.submit(tid(q), tid(q), q[2], q[2])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        lat, out_nnz, sizes = [], 0, []
        for k in range(n):
            r, info = srv.wait(ids[k])
            lat.append(info["latency_ms"])
            sizes.append(info["batch_size"])
            out_nnz += r.nnz if r is not None else 0
            del r
        wall = time.perf_counter() - t0
        return lat, wall, out_nnz, sizes

    log("server: poisson")
    lat_s, wall_s, out_s, sizes = run_server(arrivals)
    log("server: burst")
    burst_s = run_server(np.zeros(n))
    st = srv.stats()
    srv.close()

    # ---- serial plain calls --------------------------------------------------
    dsmall = [to_device(t) for t in small]
    dbig = [to_device(t) for t in big]

    def plain(q):
        if q[0] == "large":
            d = dbig[q[1]]
            return sptc.SpTC(d, d, q[2], q[2], on_device=True, count_only=True,
                             mem_budget_bytes=budget // 3)
        d = dsmall[q[1]]
        return sptc.SpTC(d, d, q[2], q[2], on_device=True)

    for q in reqs[:20]:
        del_r = plain(q)
        del del_r
    torch.cuda.synchronize()

    def run_serial(times):
        t0 = time.perf_counter()
        lat, out_nnz = [], 0
        for k, q in enumerate(reqs):
            dt = times[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            r = plain(q)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t0 - times[k]) * 1e3)
            out_nnz += r.nnz
            del r
        return lat, time.perf_counter() - t0, out_nnz

    log("serial: poisson")
    lat_p, wall_p, out_p = run_serial(arrivals)
    log("serial: burst")
    burst_p = run_serial(np.zeros(n))
    res = {
        "metric": "SpTC requests/s (served, burst)",
        "value": n / burst_s[1],
        "unit": "requests/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": burst_s[1] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serving", "requests": n, "qps": args.qps,
                   "large_fraction": float(kinds.mean()), "budget_bytes": budget},
        "server_poisson": {**_percentiles(lat_s), "wall_s": wall_s,
                           "mean_batch": float(np.mean(sizes)),
                           "small": _percentiles(np.asarray(lat_s)[~kinds]),
                           "large": _percentiles(np.asarray(lat_s)[kinds])},
        "serial_poisson": {**_percentiles(lat_p), "wall_s": wall_p,
                           "small": _percentiles(np.asarray(lat_p)[~kinds]),
                           "large": _percentiles(np.asarray(lat_p)[kinds])},
        "server_burst": {**_percentiles(burst_s[0]), "wall_s": burst_s[1],
                         "requests_per_s": n / burst_s[1],
                         "small": _percentiles(np.asarray(burst_s[0])[~kinds]),
                         "large": _percentiles(np.asarray(burst_s[0])[kinds])},
        "serial_burst": {**_percentiles(burst_p[0]), "wall_s": burst_p[1],
                         "requests_per_s": n / burst_p[1],
                         "small": _percentiles(np.asarray(burst_p[0])[~kinds]),
                         "large": _percentiles(np.asarray(burst_p[0])[kinds])},
        "burst_speedup": burst_p[1] / burst_s[1],
        "output_nnz_equal": out_s == out_p and burst_s[2] == out_p,
        "server_stats": st,
    }
    print(json.dumps(res))


def bench_nell2(args):
    """NELL-2-shaped workload: 12092x9184x28818, 7.69e7 nnz, contracted on mode 2 with
    itself under a 64 GB working-set cap.  The full contraction has 3.03e13
    intermediate products (about 9 minutes count-only), so by default the
    run contracts an evenly strided sample of X's free rows (about
    --nell2-products products) against the whole tensor, count-only, and
    extrapolates the full-run time from the measured product rate;
    --nell2-full runs the whole contraction count-only once (progress lines
    from the library on stderr) and reports its measured time."""
    import torch
    from gsparc_amd import sptc
    from gsparc_amd.coo import COOTensor
    from gsparc_amd.workloads import NELL2, NELL2_CAP_BYTES

    log("generating nell2")
    X = NELL2.make()
    if args.nell2_full:
        return nell2_full(X)
    i, j, k = (a.astype(np.int64) for a in X.idx)
    cnt = np.bincount(k, minlength=X.dims[2]).astype(np.float64)
    prod = cnt[k]
    rk = i * X.dims[1] + j
    ur, rinv = np.unique(rk, return_inverse=True)
    rowp = np.bincount(rinv, weights=prod)
    ir_total = float(rowp.sum())
    stride = max(1, int(np.ceil(ir_total / args.nell2_products)))
    pick = (np.arange(len(ur)) % stride) == 0
    sel = pick[rinv]
    Xs = COOTensor(X.dims, [a[sel] for a in X.idx], X.val[sel])
    ir_sample = float(rowp[pick].sum())
    del i, j, k, prod, rk, rinv
    dX, dS = to_device(X), to_device(Xs)
    torch.cuda.synchronize()
    phases = full_size_phases(X, [2], [2])
    log(f"full-size phases: {phases}")
    log(f"sample: every {stride}th row, {Xs.nnz} nnz, {ir_sample:.3e} products")
    times, res = [], None
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        res = sptc.SpTC(dS, dX, [2], [2], on_device=True, count_only=True,
                        mem_budget_bytes=NELL2_CAP_BYTES, kernel_timing=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"step {len(times)}: {times[-1]:.2f} s, {res.info.nchunks} chunks")
    t = float(np.median(times))
    if res.ir_nnz != int(round(ir_sample)):
        raise SystemExit(f"ir mismatch {res.ir_nnz} vs {ir_sample}")
    out = {
        "metric": "SpTC output nnz/s (nell2 row sample, count-only)",
        "value": res.nnz / t,
        "unit": "nnz/s",
        "n_gpus": 1, "steps": len(times), "warmup": 0, "ms_per_step": t * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": NELL2.name, "cap_bytes": NELL2_CAP_BYTES, "sample_stride": stride,
                   "sample_rows": int(pick.sum()), "rows": int(len(ur)),
                   "sample_nnz": Xs.nnz, "sample_ir_nnz": res.ir_nnz,
                   "sample_out_nnz": res.nnz, "chunks": res.info.nchunks},
        "products_per_s": res.ir_nnz / t,
        "full_ir_nnz": ir_total,
        "full_time_s_extrapolated": ir_total / (res.ir_nnz / t),
        "value_sum": res.info.value_sum,
        "full_size_phases": phases,
        "kernels": {k: {"launches": n, "avg_ms": ms / n, "achieved_gbs": by / (ms * 1e-3) / 1e9}
                    for k, (n, ms, by) in res.kernel_stats().items() if n and ms},
    }
    print(json.dumps(out))


def nell2_full(X):
    """The whole nell2 self-contraction, count-only, under the 64 GB cap:
    one library call, timed, with the ir closed form checked."""
    import torch
    from gsparc_amd import sptc
    from gsparc_amd.workloads import NELL2, NELL2_CAP_BYTES
    k = X.idx[2].astype(np.int64)
    cnt = np.bincount(k, minlength=X.dims[2]).astype(np.int64)
    ir_closed = int((cnt * cnt).sum())
    dX = to_device(X)
    torch.cuda.synchronize()
    log(f"full contraction: {ir_closed:.4e} products (closed form)")
    t0 = time.perf_counter()
    res = sptc.SpTC(dX, dX, [2], [2], on_device=True, count_only=True,
                    mem_budget_bytes=NELL2_CAP_BYTES, verbose=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    log(f"done: {t:.1f} s, {res.info.nchunks} chunks, {res.nnz} outputs")
    if res.ir_nnz != ir_closed:
        raise SystemExit(f"ir mismatch {res.ir_nnz} vs {ir_closed}")
    print(json.dumps({
        "metric": "SpTC output nnz/s (nell2 full, count-only)",
        "value": res.nnz / t, "unit": "nnz/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": NELL2.name, "cap_bytes": NELL2_CAP_BYTES, "nnz": X.nnz,
                   "ir_nnz": res.ir_nnz, "ir_closed_form": ir_closed, "out_nnz": res.nnz,
                   "chunks": res.info.nchunks},
        "products_per_s": res.ir_nnz / t,
        "value_sum": res.info.value_sum,
    }))


def full_size_phases(X, cx, cy, reps=5):
    """Operand sort and index match over the whole operand through the
    per-phase entry points (sptc_linearize, sptc_sort_pairs,
    sptc_index_match), timed with HIP events on the launch stream: median
    ms and algorithmic GB/s (sort: 24 B per key per radix pass, key + value
    read and written; match: 8 B key read, 24 B of mpos/mcnt/prefix written,
    per X nonzero, plus 8 B per Y key for the contract table)."""
    import ctypes
    import torch
    from gsparc_amd import sptc
    import oracle as orc
    L = sptc.lib()
    cw = orc.common_cwidth(X.dims, X.dims, cx, cy)
    res = {}
    for role, cm in ((1, cx), (2, cy)):
        sh, wd, nf, ncb = sptc.layout(X.dims, cm, role, cw)
        idx = [torch.from_numpy(a.view(np.int32)).cuda() for a in X.idx]
        ptrs = (ctypes.c_void_p * X.order)(*[a.data_ptr() for a in idx])
        keys = torch.empty(X.nnz, dtype=torch.int64, device="cuda")
        vals = torch.from_numpy(X.val).cuda()
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        sptc.check(L.sptc_linearize(X.order, X.nnz, ptrs, 4, (ctypes.c_int32 * X.order)(*sh),
                                    (ctypes.c_int32 * X.order)(*wd),
                                    ctypes.c_void_p(keys.data_ptr()), stream))
        base_k, base_v = keys.clone(), vals.clone()
        nbits = nf + ncb
        t = []
        for _ in range(reps):
            keys.copy_(base_k)
            vals.copy_(base_v)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            sptc.check(L.sptc_sort_pairs(ctypes.c_void_p(keys.data_ptr()),
                                         ctypes.c_void_p(vals.data_ptr()), X.nnz, nbits, 0,
                                         stream))
            e1.record()
            torch.cuda.synchronize()
            t.append(e0.elapsed_time(e1))
        ms = float(np.median(t))
        res[f"sort_{'x' if role == 1 else 'y'}"] = {
            "ms": ms, "passes_assumed_8bit": -(-nbits // 8),
            "gbs_per_8bit_pass_model": 24.0 * X.nnz * (-(-nbits // 8)) / (ms * 1e-3) / 1e9}
        if role == 1:
            xk, ncbits = keys, ncb
        else:
            yk, ynf = keys, nf
    nx = X.nnz
    mpos = torch.empty(nx, dtype=torch.int64, device="cuda")
    mcnt = torch.empty(nx, dtype=torch.int64, device="cuda")
    pre = torch.empty(nx + 1, dtype=torch.int64, device="cuda")
    ir = ctypes.c_uint64()
    t = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        sptc.check(L.sptc_index_match(ctypes.c_void_p(xk.data_ptr()), nx,
                                      ctypes.c_void_p(yk.data_ptr()), nx, ncbits, ynf,
                                      ctypes.c_void_p(mpos.data_ptr()),
                                      ctypes.c_void_p(mcnt.data_ptr()),
                                      ctypes.c_void_p(pre.data_ptr()), ctypes.byref(ir), stream))
        e1.record()
        torch.cuda.synchronize()
        t.append(e0.elapsed_time(e1))
    ms = float(np.median(t))
    res["match"] = {"ms": ms, "ir": ir.value,
                    "gbs": (32.0 * nx + 8.0 * nx) / (ms * 1e-3) / 1e9}
    return res


def self_check(X, spec, r):
    """Checks the last timed step's result against what the input alone
    fixes: ir = sum over X nonzeros of the count of nonzeros sharing their
    contracted coordinates, the FP64 checksum of the products =
    sum_c S(c)^2 (S(c) = sum of values at contracted coordinates c), and one
    device pass over every emitted entry (FP64 sum of the values, strict
    (row, column) order).  Exits non-zero on a mismatch."""
    ck = np.zeros(X.nnz, np.int64)
    for m in spec.cx:
        ck = ck * np.int64(X.dims[m]) + X.idx[m].astype(np.int64)
    u, inv = np.unique(ck, return_inverse=True)
    cnt = np.bincount(inv).astype(np.int64)
    S = np.bincount(inv, weights=X.val.astype(np.float64))
    ir_cf = int((cnt * cnt).sum())
    vs_cf = float((S * S).sum())
    emitted, bad = r.verify()
    u32 = 2.0 ** -24
    out = {"ir_nnz": r.ir_nnz, "ir_closed_form": ir_cf, "value_sum": r.info.value_sum,
           "value_sum_closed_form": vs_cf, "emitted_value_sum": emitted,
           "order_violations": bad}
    # positive values: FP32-rounded products sum within u32 of the exact
    # checksum, compensated FP32 output sums within about 3 u32
    ok = (r.ir_nnz == ir_cf and abs(r.info.value_sum - vs_cf) <= 2 * u32 * vs_cf
          and abs(emitted - vs_cf) <= 6 * u32 * vs_cf and bad == 0)
    out["ok"] = bool(ok)
    log(f"self-check: {out}")
    if not ok:
        print(json.dumps({"error": "result check failed", "check": out}))
        raise SystemExit(3)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: the library "
              f"runs on one device, and sharding across GPUs is a design, not code",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "batch_stated":
        return bench_batch_stated(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "nell2":
        return bench_nell2(args)
    import torch
    from gsparc_amd import sptc
    from gsparc_amd.workloads import CONFIGS, NELL2_CAP_BYTES

    spec = CONFIGS[args.workload]
    X = spec.make()
    dX = to_device(X)
    torch.cuda.synchronize()
    kw = {}
    if args.workload == "nell2":
        kw = dict(count_only=True, mem_budget_bytes=NELL2_CAP_BYTES)

    if args.copy_probe_bytes:
        a = torch.empty(args.copy_probe_bytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    def step(timing):
        return sptc.SpTC(dX, dX, list(spec.cx), list(spec.cy), on_device=True,
                         radix_bits=args.radix_bits, kernel_timing=timing,
                         dense_accumulator=args.accumulator, **kw)

    for _ in range(args.warmup):
        r = step(False)
        del r
    torch.cuda.synchronize()
    times, phases = [], []
    kstats = {}
    out_nnz = ir = acc = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    last = None
    for k_step in range(args.steps):
        ts = time.perf_counter()
        r = step(True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - ts)
        phases.append(r.phase_ms)
        for k, (n, kms, kb) in r.kernel_stats().items():
            a0 = kstats.get(k, (0, 0.0, 0.0))
            kstats[k] = (a0[0] + n, a0[1] + kms, a0[2] + kb)
        out_nnz, ir, acc = r.nnz, r.ir_nnz, r.info.accumulator
        if k_step == args.steps - 1:
            last = r  # checked after the timed region
        del r
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    check = self_check(X, spec, last)
    del last
    ph = np.median(np.array(phases), axis=0).tolist()
    res = {
        "metric": "SpTC output nnz/s",
        "value": out_nnz / (elapsed / args.steps),
        "unit": "nnz/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": spec.name, "contract_x": list(spec.cx),
                   "contract_y": list(spec.cy), "nnz": X.nnz, "ir_nnz": ir,
                   "out_nnz": out_nnz, "accumulator": ["esc", "dense", "tiled_dense"][acc]},
        "phase_ms_median": {"linearize_sort": ph[1], "match": ph[2], "partition": ph[3],
                            "accumulate": ph[4], "total": ph[5]},
        "check": check,
    }
    peak = sptc.device_info()["peak_hbm_gbs"]
    kern = {}
    for k, (n, kms, kb) in kstats.items():
        if n:
            kern[k] = {"launches": n, "avg_ms": kms / n, "alg_bytes_per_launch": kb / n,
                       "achieved_gbs": kb / (kms * 1e-3) / 1e9 if kms else None}
    res["kernels"] = kern
    if kern:
        dom = max(kern, key=lambda k: kstats[k][1])
        d = kern[dom]
        res["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": d["achieved_gbs"],
                           "peak": peak, "unit": "GB/s", "frac": d["achieved_gbs"] / peak,
                           "traffic": measured_traffic(args.workload, dom),
                           "avg_launch_ms": d["avg_ms"],
                           "alg_bytes_per_launch": d["alg_bytes_per_launch"]}
    res["pool"] = sptc.pool_stats()
    if not args.no_cpu_baseline and args.workload != "nell2":
        target = args.cpu_sample_products or (7e8 if args.workload == "nips" else 8e7)
        cb = cpu_baseline(spec, X, target)
        if cb:
            res["cpu_baseline"] = {
                "value": cb["sample_out_nnz"] / cb["total_s"], "unit": "nnz/s",
                "cores": cb["threads"], "kind": "reference", "cpu_model": cb["cpu_model"],
                "sample": (f"Sparta ttt (EXPERIMENT_MODES=3, [Total time]) on every "
                           f"{cb['stride']}th of the {cb['rows']} X free rows "
                           f"({cb['sample_rows']} rows, {cb['sample_nnz']} nnz, "
                           f"{cb['sample_products']} products, {cb['sample_out_nnz']} output "
                           f"nnz) against all of X: {cb['total_s']:.3f} s"),
                "products_per_s": cb["sample_products"] / cb["total_s"],
                "gpu_products_per_s": ir / (elapsed / args.steps),
            }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
