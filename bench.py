#!/usr/bin/env python3
"""Benchmark of the graph-ANN query path on one MI355X.

Workload (default): the 1M x 128 fp32 Gaussian mixture (1,000 components,
sigma 0.25, seeded), NSG-style graph with R=32 built on the GPU, 10,000
queries, L2, K=10.  The L sweep {32, 64, 128, 256} picks the smallest L
whose recall@10 against exact ground truth reaches 0.95; at that L:

  value        whole-batch QPS of the batched persistent kernel (one
               hipGraph-replayed launch over all queries per step, inputs
               resident in HBM), timed over --steps steps;
  plain_qps    the same queries through one kernel launch per query;
  latency      p50/p99 of per-query device service time in the batch;
  roofline     algorithmic bytes per launch (sum over queries of
               NDC*dim*4 + hops*R*4 + dim*4 + K*8) / kernel time vs 8 TB/s;
  cpu_baseline the ScatterSearch intra-query search (CPU restatement) on
               the host cores over a bounded sample of the same queries.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC_NAME = "QPS at recall@10=0.95, batched vs plain kernel, 1 MI355X; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def host_cpus():
    """(threads the CPU baseline uses, CPUs this process may run on, CPUs the
    OS reports).  The baseline uses every CPU it may run on, at most 16: a
    one-GPU box's share of its host is 16 CPUs, whatever nproc reports."""
    n_os = os.cpu_count() or 1
    try:
        n_aff = len(os.sched_getaffinity(0))
    except AttributeError:
        n_aff = n_os
    return max(1, min(n_aff, 16)), n_aff, n_os


def host_threads():
    return host_cpus()[0]


def cpu_model():
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                info[k] = v
            if k == "flags":
                break
    except OSError:
        pass
    name = info.get("model name", "unknown")
    return (f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}, {info.get('cpu MHz', '?')} MHz)")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--num", type=int, default=0, help="override the base size")
    ap.add_argument("--nq", type=int, default=0, help="override the query count")
    ap.add_argument("--L", type=int, default=0, help="fixed L instead of the sweep")
    ap.add_argument("--recall-target", type=float, default=0.95)
    ap.add_argument("--plain-queries", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the wider-expansion and several-wavefronts-per-query measurements")
    ap.add_argument("--visited-slots", type=int, default=0)
    ap.add_argument("--waves-per-block", type=int, default=0)
    ap.add_argument("--visited-bucket", type=int, default=0, help="16-bit visited-table bucket (8 or 16); 0 = tuned")
    ap.add_argument("--team-size", type=int, default=0, help="lanes per distance (4 or 8); 0 = tuned")
    ap.add_argument("--pool-layout", type=int, default=0, help="SS_POOL_* (1 registers, 2 LDS); 0 = tuned")
    ap.add_argument("--pool-classes", choices=["auto", "off", "serial"], default="auto",
                    help="--profile of the mixed launch: the class kernels side by side (auto), one kernel "
                         "longest pool first (off), or the class kernels one after another (serial)")
    ap.add_argument("--query-order", choices=["auto", "index", "locality"], default="auto",
                    help="order in which the batched search takes the queries (results do not depend on it)")
    ap.add_argument("--graph-cache", default="",
                    help="NSG file: load the graph from it if it exists, else build and write it")
    ap.add_argument("--profile", action="store_true",
                    help="only build, then run --steps batched launches (for rocprofv3)")
    ap.add_argument("--serve-rates", default="50000,100000,200000,300000,500000,800000,1000000",
                    help="Poisson arrival rates (QPS) of the serving sweep; empty to skip")
    ap.add_argument("--batch-sizes", default="1,64,1024,10000",
                    help="also time hipGraph-replayed launches of these batch sizes; empty to skip")
    ap.add_argument("--large-L", default="1024,1600",
                    help="also time batched launches at these pools above 512 entries; empty to skip")
    ap.add_argument("--serve-config", default="",
                    help='serving knobs as JSON, e.g. {"standby_waves": 1024} (ss_serve_config fields)')
    ap.add_argument("--serve-seconds", type=float, default=0.5,
                    help="length of the arrival stream at each serving rate (at least 20,000 requests)")
    ap.add_argument("--gt-queries", type=int, default=-1,
                    help="queries with exact ground truth (recall is measured on these); -1 = all, "
                         "or the first 10,000 on the 10M workload")
    args = ap.parse_args()
    if args.workload == "c3":
        # the 10M x 96 workload: L in {64, 128} as configured, batch sizes up to the
        # whole 100,000-query batch; its graph is not the serving workload's
        if args.gt_queries < 0:
            args.gt_queries = 10000
        if "--serve-rates" not in sys.argv:
            args.serve_rates = ""
        if "--batch-sizes" not in sys.argv:
            args.batch_sizes = "1,64,1024,10000,100000"
        args.no_variants = True
        if "--large-L" not in sys.argv:
            args.large_L = ""

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the query path "
              "shards by query across GPUs with no exchange, so N GPUs run N independent copies)",
              file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401  (shares its HIP runtime with the library)
    import scattersearch_amd as ss
    from scattersearch_amd import data, metrics
    from scattersearch_amd import io as ssio
    from scattersearch_amd.graph import gpu_build_graph

    torch.cuda.set_device(0)
    w = data.WORKLOADS[args.workload]
    num = args.num or w.num
    nq = args.nq or w.nq
    t0 = time.time()
    base, queries = data.generate(w, num, nq)
    log(f"data {num}x{w.dim} + {nq} queries in {time.time() - t0:.1f}s")

    ix = ss.Index(base, w.metric)
    t0 = time.time()
    if args.graph_cache and os.path.exists(args.graph_cache):
        adj, ep = ssio.read_nsg(args.graph_cache, r_pad=w.R)
        ix.set_graph(adj, ep)
        ginfo = {"passes": [{"mean_degree": float((adj != 0xFFFFFFFF).sum(1).mean())}], "knn_s": 0.0,
                 "cached": args.graph_cache}
        log(f"graph read from {args.graph_cache}, ep {ep}")
    else:
        adj, ep, ginfo = gpu_build_graph(ix, base, w.R, w.knn_k, w.nav_L, log=log,
                                         nn_descent_iters=w.extra.get("nn_descent_iters", 0),
                                         knn_chunk=131072 if num * w.dim > 150_000_000 else 0)
        log(f"graph built in {time.time() - t0:.1f}s, ep {ep}")
        if args.graph_cache:
            ssio.write_nsg(args.graph_cache, adj, ep)
    r_pad = ix.info()["r_pad"]
    K = w.K
    if not w.L_list:  # the uneven-queries workload: one launch mixing K and L
        return mixed_run(args, w, ix, queries, ginfo, r_pad)

    dim_pad = ss.dim_padded(w.dim)
    dq = torch.zeros((nq, dim_pad), dtype=torch.float32, device="cuda")
    dq[:, :w.dim] = torch.from_numpy(queries).cuda()
    d_ids = torch.empty((nq, K), dtype=torch.int32, device="cuda")
    d_dist = torch.empty((nq, K), dtype=torch.float32, device="cuda")
    d_stats = torch.empty((nq, 8), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()  # default priority; hipGraph capture needs a non-null stream

    qorder = {"auto": ss._lib.SS_ORDER_AUTO, "index": ss._lib.SS_ORDER_INDEX,
              "locality": ss._lib.SS_ORDER_LOCALITY}[args.query_order]

    def order_used(n, wq=1):
        if qorder == ss._lib.SS_ORDER_AUTO:
            return "locality" if n >= ss._lib.SS_ORDER_MIN_BATCH and wq <= 1 else "index"
        return args.query_order

    def launch(L, use_graph=True, slots=args.visited_slots, wpb=args.waves_per_block, ew=0, wq=0,
               team=args.team_size, layout=args.pool_layout):
        ix.search_device(dq.data_ptr(), nq, K, L, K, d_ids.data_ptr(), d_dist.data_ptr(),
                         d_stats.data_ptr(), stream=stream.cuda_stream, use_graph=use_graph,
                         visited_slots=slots, waves_per_block=wpb, expand_width=ew, waves_per_query=wq,
                         query_order=qorder, team_size=team if ew in (0, 1) and wq in (0, 1) else 0,
                         pool_layout=layout if ew in (0, 1) and wq in (0, 1) else 0,
                         visited_bucket=args.visited_bucket if ew in (0, 1) and wq in (0, 1) else 0)

    def stats_host():
        st = d_stats.cpu().numpy().view(np.uint32).reshape(-1).view(ss.STATS_DTYPE)
        if (st["status"] == ss._lib.SS_QUERY_FAULT).any():
            raise RuntimeError("search reported an internal error (node id outside the index)")
        return st

    if args.profile:
        L = args.L or 64
        with torch.cuda.stream(stream):
            for _ in range(args.steps):
                launch(L)
            # calibration for HBM counters: a plain copy of a known size
            # (512 MiB read + 512 MiB written per launch)
            src = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(5):
                dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile: {args.steps} launches at L={L} and 5 copies of 512 MiB done")
        return

    t0 = time.time()
    n_gt = nq if args.gt_queries < 0 else min(nq, args.gt_queries)
    gt, _ = ix.bruteforce(queries[:n_gt], K)
    log(f"ground truth of {n_gt} queries in {time.time() - t0:.1f}s")

    def recall_of(ids):
        return metrics.recall(ids[:n_gt], gt, K)

    def timed(L, steps, warmup, **kw):
        for _ in range(warmup):
            launch(L, **kw)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t_wall = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            launch(L, **kw)
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall
        kern_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
        return wall, kern_ms

    sweep = []
    L_list = [args.L] if args.L else list(w.L_list)
    chosen = None
    for L in L_list:
        wall, kern_ms = timed(L, 3, 1)
        ids = d_ids.cpu().numpy().view(np.uint32)
        st = stats_host()
        rec = recall_of(ids)
        row = dict(L=L, recall=round(rec, 4), qps=round(nq * 3 / wall, 1), kernel_ms=round(kern_ms, 3),
                   mean_ndc=round(float(st["ndc"].mean()), 1), mean_hops=round(float(st["hops"].mean()), 1),
                   lossy=int(st["lossy"].sum()))
        sweep.append(row)
        log(f"sweep {row}")
        if chosen is None and rec >= args.recall_target:
            chosen = L
    if chosen is None:
        chosen = L_list[-1]
        log(f"no L reached recall {args.recall_target}; reporting the largest L")
    elif not args.L and args.workload == "c1":
        # the smallest L that reaches the target: refine between the last L
        # below it and the first above it in steps of 8, 2, then 1
        lo = max([r["L"] for r in sweep if r["L"] < chosen] or [K])
        for step in (8, 2, 1):
            for Lr in range(lo + step, chosen, step):
                wall, kern_ms = timed(Lr, 3, 1)
                ids = d_ids.cpu().numpy().view(np.uint32)
                rec = recall_of(ids)
                st = stats_host()
                row = dict(L=Lr, recall=round(rec, 4), qps=round(nq * 3 / wall, 1), kernel_ms=round(kern_ms, 3),
                           mean_ndc=round(float(st["ndc"].mean()), 1),
                           mean_hops=round(float(st["hops"].mean()), 1), lossy=int(st["lossy"].sum()))
                sweep.append(row)
                log(f"refine {row}")
                if rec >= args.recall_target:
                    chosen = Lr
                    break
                lo = Lr
    L = chosen
    sweep.sort(key=lambda r: r["L"])

    # Autotune launch parameters for this L on a sample (exact results kept).
    tuned = ix.autotune(queries, K, L, max_recall_loss=0.0)
    log(f"autotuned for L={L}: {tuned}")

    # Timed region: K steps of one batched launch over all queries.
    wall, kern_ms = timed(L, args.steps, args.warmup)
    qps = nq * args.steps / wall
    ms_per_step = wall * 1e3 / args.steps
    ids = d_ids.cpu().numpy().view(np.uint32)
    st = stats_host()
    rec = recall_of(ids)
    khz = ix.wall_clock_khz()
    service_ms = (st["t_end"].astype(np.float64) - st["t_start"]) / khz
    done_ms = (st["t_end"].astype(np.float64) - st["t_start"].min()) / khz
    ndc_exact = exact_ndc(ids, st, lambda slots: launch(L, use_graph=False, slots=slots), d_ids, stats_host)
    bq = metrics.bytes_per_query(ndc_exact, st["hops"], w.dim, r_pad, K)
    launch_bytes = float(bq.sum())
    achieved_gbs = launch_bytes / (kern_ms * 1e-3) / 1e9
    log(f"batched L={L}: {qps:.0f} QPS, kernel {kern_ms:.3f} ms, {achieved_gbs:.0f} GB/s algorithmic, "
        f"recall {rec:.4f}; distances recomputed after the visited hash forgot an id: "
        f"{float((st['ndc'] - ndc_exact).mean()):.1f} per query (not counted)")

    # The same batch through host buffers (ss_search: queries copied in and
    # results copied back over PCIe around the launch), for reference.
    ix.search(queries, K, L)
    t_h = time.perf_counter()
    for _ in range(3):
        ix.search(queries, K, L)
    host_qps = nq * 3 / (time.perf_counter() - t_h)
    log(f"host-buffer batches (PCIe copies included): {host_qps:.0f} QPS")

    # Plain kernel: one launch per query.
    npq = min(args.plain_queries, nq)
    from scattersearch_amd import SS_MODE_PER_QUERY

    def plain(n):
        ix.search_device(dq.data_ptr(), n, K, L, K, d_ids.data_ptr(), d_dist.data_ptr(),
                         d_stats.data_ptr(), stream=stream.cuda_stream, mode=SS_MODE_PER_QUERY,
                         visited_slots=args.visited_slots)

    plain(min(64, npq))
    torch.cuda.synchronize()
    t = time.perf_counter()
    plain(npq)
    torch.cuda.synchronize()
    plain_qps = npq / (time.perf_counter() - t)
    log(f"plain per-query launches: {plain_qps:.0f} QPS over {npq} queries")

    # Launches of several batch sizes (persistent, hipGraph-replayed), at L.
    batches = []
    for bsz in [int(x) for x in args.batch_sizes.split(",") if x]:
        bsz = min(bsz, nq)

        def launch_b(n=bsz):
            ix.search_device(dq.data_ptr(), n, K, L, K, d_ids.data_ptr(), d_dist.data_ptr(),
                             d_stats.data_ptr(), stream=stream.cuda_stream, use_graph=True,
                             query_order=qorder)

        reps = max(3, min(200, 20000 // bsz))
        launch_b()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        t_b = time.perf_counter()
        for a_ev, b_ev in evs:
            a_ev.record(stream)
            launch_b()
            b_ev.record(stream)
        torch.cuda.synchronize()
        wall_b = time.perf_counter() - t_b
        kms = [a_ev.elapsed_time(b_ev) for a_ev, b_ev in evs]
        row = dict(batch=bsz, launches=reps, qps=round(bsz * reps / wall_b, 1),
                   launch_ms_p50=round(float(np.percentile(kms, 50)), 4),
                   launch_ms_p99=round(float(np.percentile(kms, 99)), 4), query_order=order_used(bsz))
        batches.append(row)
        log(f"batch {row}")

    # Relaxed variants, judged by recall: wider expansion (ModifiedDeltaStepping)
    # and several wavefronts per query (ScatterSearch).
    variants = []
    if not args.no_variants:
        for ew, wq in ((2, 1), (4, 1), (1, 2), (1, 4)):
            rows = []
            for Lv in sorted(set([32, 40, 48, 64, L])):
                slots = 8192 if wq > 1 else 0
                try:
                    wall_v, kern_v = timed(Lv, 3, 1, slots=slots, ew=ew, wq=wq)
                except Exception as exc:  # a shape that does not fit is reported, not fatal
                    rows.append(dict(L=Lv, error=str(exc)))
                    continue
                st_v = stats_host()
                rows.append(dict(L=Lv, recall=round(recall_of(d_ids.cpu().numpy().view(np.uint32)), 4),
                                 qps=round(nq * 3 / wall_v, 1), kernel_ms=round(kern_v, 3),
                                 mean_ndc=round(float(st_v["ndc"].mean()), 1)))
            ok = [r for r in rows if r.get("recall", 0) >= args.recall_target]
            best_v = max(ok, key=lambda r: r["qps"]) if ok else None
            variants.append(dict(expand_width=ew, waves_per_query=wq, best_at_target=best_v, sweep=rows))
            log(f"variant ew={ew} wq={wq}: best at recall>={args.recall_target}: {best_v}")

    # Pools above 512 entries (the large-pool kernel; the reference's own
    # sweeps run L up to 1,600 and 8,000-16,000).
    large = []
    for Lg in [int(x) for x in args.large_L.split(",") if x]:
        wall_g, kern_g = timed(Lg, 2, 1)
        st_g = stats_host()
        large.append(dict(L=Lg, recall=round(recall_of(d_ids.cpu().numpy().view(np.uint32)), 4),
                          qps=round(nq * 2 / wall_g, 1), kernel_ms=round(kern_g, 3),
                          mean_ndc=round(float(st_g["ndc"].mean()), 1),
                          mean_hops=round(float(st_g["hops"].mean()), 1)))
        log(f"large pool {large[-1]}")

    serving = None
    if args.serve_rates:
        try:
            serving = serve_sweep(ix, queries, gt, w, [float(r) for r in args.serve_rates.split(",")],
                                  args.serve_seconds, json.loads(args.serve_config) if args.serve_config else None)
        except Exception as exc:  # the batch numbers above stay valid
            serving = {"error": str(exc)}
            log(f"serving sweep failed: {exc}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(base, adj, ep, queries, gt, K, L, w.metric, args.cpu_seconds)

    out = {
        "metric": METRIC_NAME,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: seeded {w.centres}-component Gaussian mixture (seed {w.seed}), graph built on the GPU",
        "config": {"workload": w.name, "num": num, "dim": w.dim, "nq": nq, "R": w.R, "K": K, "L": L,
                   "metric": "L2" if w.metric == 0 else "inner_product", "recall@10": round(rec, 4),
                   "mean_ndc": round(float(ndc_exact.mean()), 1), "mean_hops": round(float(st["hops"].mean()), 1),
                   "recomputed_ndc": round(float((st["ndc"] - ndc_exact).mean()), 1),
                   "launch": dict({k: tuned[k] for k in ("expand_width", "waves_per_block", "visited_slots",
                                                         "visited_bucket", "team_size", "pool_layout")},
                                  query_order=order_used(nq))},
        "host_buffers_qps": round(host_qps, 1),
        "plain_qps": round(plain_qps, 1),
        "speedup_vs_plain": round(qps / plain_qps, 2),
        "serving": serving,
        "variants": variants,
        "large_pools": large,
        "batches": batches,
        "latency_ms": {"service_p50": round(metrics.percentile(service_ms, 0.5), 4),
                       "service_p99": round(metrics.percentile(service_ms, 0.99), 4),
                       "batch_completion_p50": round(metrics.percentile(done_ms, 0.5), 4),
                       "batch_completion_p99": round(metrics.percentile(done_ms, 0.99), 4)},
        "l_sweep": sweep,
        "graph": {"mean_degree": ginfo["passes"][-1]["mean_degree"], "knn_s": round(ginfo["knn_s"], 1),
                  "passes": len(ginfo["passes"])},
        "roofline": roofline_record(w.name, nq, K, L, order_used(nq), achieved_gbs, launch_bytes, kern_ms,
                                    match={"visited_slots": args.visited_slots or tuned["visited_slots"],
                                           "visited_bucket": args.visited_bucket or tuned["visited_bucket"]}),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def exact_ndc(ids, st, launch_with_slots, d_ids, stats_host):
    """Distance computations of the exact search for the timed launch's
    queries: the same launch once more (untimed) with a visited hash large
    enough to forget nothing.  A forgotten id only makes the search compute
    its distance again (same ids, same hops), so these counts, not the timed
    launch's, are the algorithmic ones."""
    import scattersearch_amd as ss
    best = st["ndc"].astype(np.int64)
    for slots in (16384, 32768):
        try:
            launch_with_slots(slots)
        except ss.SSError:
            continue
        import torch
        torch.cuda.synchronize()
        st2 = stats_host()
        if not np.array_equal(d_ids.cpu().numpy().view(np.uint32), ids) or \
                not np.array_equal(st2["hops"], st["hops"]):
            raise RuntimeError("the search with a larger visited hash returned other results")
        best = np.minimum(best, st2["ndc"].astype(np.int64))
        if st2["lossy"].sum() == 0:
            return st2["ndc"].astype(np.int64)
        log(f"visited hash of {slots} slots: {int((st2['lossy'] > 0).sum())} queries still forgot ids")
    log("counting, for queries whose visited hash forgot ids, the fewest distance computations seen")
    return best


def measured_traffic(workload, nq, K, L, order, match=None):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    calibrated against a plain copy) for this workload and launch, from the
    records kept in profiles/search_kernel_<workload>_counters.md (one JSON
    block per launch shape measured): the record of exactly this launch
    (workload, batch, K, L, query order and the fields of `match`), else one
    of the same workload, batch, K, L and order measured at another shape
    (returned with "shape_differs" set), else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"search_kernel_{workload}_counters.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    near = None
    for block in text.split("```json")[1:]:
        try:
            rec = json.loads(block.split("```", 1)[0])
        except ValueError:
            continue
        if (rec.get("workload"), rec.get("nq"), rec.get("K"), rec.get("L"), rec.get("query_order", "index")) != \
                (workload, nq, K, L, order):
            continue
        if any(rec.get(k) != v for k, v in (match or {}).items()):
            near = near or dict(rec, shape_differs=True)
            continue
        return rec
    return near


def roofline_record(workload, nq, K, L, order, achieved_gbs, launch_bytes, kern_ms, match=None):
    rec = measured_traffic(workload, nq, K, L, order, match)
    out = {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
           "traffic": rec["traffic_bytes"] if rec else None,
           "bytes_per_launch": round(launch_bytes), "kernel_ms": round(kern_ms, 4)}
    if rec:
        out["traffic_note"] = (f"rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (copy-calibrated), "
                               f"profiles/search_kernel_{workload}_counters.md; L2 hit rate {rec['l2_hit_rate']}")
        if rec.get("shape_differs"):
            shape = {k: rec.get(k) for k in (match or {})}
            out["traffic_note"] += f"; measured at {shape}, this launch {match}"
    return out


def serve_sweep(ix, queries, gt10, w, rates, seconds, cfg=None):
    """Online serving (the C4 workload): Poisson arrivals at each rate into
    the serving loop for `seconds` (at least 20,000 requests), K = 10 (80%)
    or 100 (20%), L uniform in max(64, K)..512; latency = completion seen by
    the host minus arrival.  cfg: ss_serve_config fields to set."""
    from scattersearch_amd import metrics, serve
    out = []
    for r in rates:
        nreq = max(20000, int(r * seconds))
        rng = np.random.default_rng(20261207)
        qi = np.arange(nreq) % queries.shape[0]
        K = np.where(rng.random(nreq) < 0.8, 10, 100).astype(np.uint32)
        L = rng.integers(np.maximum(64, K), 513).astype(np.uint32)  # K <= L
        arr = serve.poisson_arrivals(nreq, r, seed=int(r))
        t = time.perf_counter()
        ids, ds, lat, st, tr = serve.replay(ix, queries[qi], K, L, arr, k_max=100, l_max=512, with_trace=True,
                                            **(cfg or {}))
        wall = time.perf_counter() - t
        span_s = (arr[-1] + lat[-1]) * 1e-6
        khz = ix.wall_clock_khz()
        svc = (st["t_end"].astype(np.float64) - st["t_start"]) / khz
        rec = metrics.recall(ids[:, :10], gt10[qi], 10)
        row = {"rate_qps": r, "requests": nreq, "achieved_qps": round(nreq / max(span_s, 1e-9), 1),
               "p50_ms": round(metrics.percentile(lat, 0.5) / 1e3, 4),
               "p99_ms": round(metrics.percentile(lat, 0.99) / 1e3, 4),
               "mean_ms": round(float(lat.mean()) / 1e3, 4),
               "service_p50_ms": round(metrics.percentile(svc, 0.5), 4),
               "service_p99_ms": round(metrics.percentile(svc, 0.99), 4),
               "recall@10": round(rec, 4), "mean_ndc": round(float(st["ndc"].mean()), 1), "wall_s": round(wall, 2)}
        # where the time went (host clock; device times mapped by a clock reading)
        parts = {"submit_lag": tr["submit_us"] - arr, "claim_wait": tr["claim_us"] - tr["submit_us"],
                 "notify": tr["seen_us"] - tr["finish_us"]}
        row["split_ms"] = {k: [round(metrics.percentile(v, 0.5) / 1e3, 4), round(metrics.percentile(v, 0.99) / 1e3, 4)]
                           for k, v in parts.items()}
        busy = tr["waves_alive"].astype(np.int64) - tr["waves_standby"]
        row["waves"] = {"alive_p50": int(np.median(tr["waves_alive"])), "alive_min": int(tr["waves_alive"].min()),
                        "busy_p50": int(np.median(busy)), "busy_max": int(busy.max())}
        out.append(row)
        log(f"serve {row}")
    return {"workload": "serving kernel with continuous admission (wavefronts claim requests from a host ring "
                        "as they arrive), Poisson arrivals, K=10 (80%) / 100 (20%), L uniform max(64, K)..512",
            "rates": out}


def mixed_run(args, w, ix, queries, ginfo, r_pad):
    """One persistent launch over all queries with K drawn from {1, 10, 100}
    and L uniform in K..512 per query (seeded), launch parameters autotuned
    at K=100, L=512 with no recall loss (exact search).  Prints one JSON line:
    QPS, recall@K per K group against exact ground truth, the algorithmic
    bytes rate against the HBM peak."""
    import torch
    import scattersearch_amd as ss
    from scattersearch_amd import metrics
    nq = queries.shape[0]
    rng = np.random.default_rng(w.seed + 17)
    Ks = rng.choice(np.array([1, 10, 100], dtype=np.uint32), size=nq)
    Ls = np.array([rng.integers(int(k), 513) for k in Ks], dtype=np.uint32)
    k_max, l_max = 100, 512
    if not args.profile:
        t0 = time.time()
        gt, _ = ix.bruteforce(queries, k_max)
        log(f"ground truth (top {k_max}) in {time.time() - t0:.1f}s")
    # the launch splits the queries by pool size (L <= 64, <= 128, <= 256,
    # <= 512); each class runs with the parameters tuned for its bound
    tuned_cls = {}
    for bound in (64, 128, 256, 512):
        tuned_cls[bound] = ix.autotune(queries, min(k_max, 10 if bound == 64 else k_max), bound,
                                       max_recall_loss=0.0)
        log(f"autotuned for L<={bound}: {tuned_cls[bound]}")
    tuned = tuned_cls[l_max]
    dim_pad = ss.dim_padded(w.dim)
    dq = torch.zeros((nq, dim_pad), dtype=torch.float32, device="cuda")
    dq[:, :w.dim] = torch.from_numpy(queries).cuda()
    dK = torch.from_numpy(Ks.astype(np.int32)).cuda()
    dL = torch.from_numpy(Ls.astype(np.int32)).cuda()
    d_ids = torch.empty((nq, k_max), dtype=torch.int32, device="cuda")
    d_dist = torch.empty((nq, k_max), dtype=torch.float32, device="cuda")
    d_stats = torch.empty((nq, 8), dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def launch(classes, slots=args.visited_slots, bucket=0):
        ix.search_device(dq.data_ptr(), nq, k_max, l_max, k_max, d_ids.data_ptr(), d_dist.data_ptr(),
                         d_stats.data_ptr(), d_K=dK.data_ptr(), d_L=dL.data_ptr(), stream=stream.cuda_stream,
                         use_graph=True, visited_slots=slots, waves_per_block=args.waves_per_block,
                         pool_classes=classes, visited_bucket=bucket)

    def timed(classes, slots=args.visited_slots, bucket=0):
        for _ in range(args.warmup):
            launch(classes, slots, bucket)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        t_wall = time.perf_counter()
        for e0, e1 in evs:
            e0.record(stream)
            launch(classes, slots, bucket)
            e1.record(stream)
        torch.cuda.synchronize()
        wall_ = time.perf_counter() - t_wall
        ids_ = d_ids.cpu().numpy().view(np.uint32)
        st_ = d_stats.cpu().numpy().view(np.uint32).reshape(-1).view(ss.STATS_DTYPE)
        if (st_["status"] != ss._lib.SS_QUERY_OK).any():
            raise RuntimeError("a query of the mixed launch did not complete")
        return wall_, float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs])), ids_, st_

    if args.profile:
        from scattersearch_amd._lib import SS_CLASSES_AUTO, SS_CLASSES_OFF, SS_CLASSES_SERIAL
        mode = {"auto": SS_CLASSES_AUTO, "off": SS_CLASSES_OFF, "serial": SS_CLASSES_SERIAL}[args.pool_classes]
        with torch.cuda.stream(stream):
            for _ in range(args.steps):
                launch(mode, args.visited_slots, args.visited_bucket)
            src = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(5):
                dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile: {args.steps} mixed launches and 5 copies of 512 MiB done")
        return

    # The class kernels side by side (the library's default), one after
    # another, and one kernel for every pool size (longest pool first), each
    # with the per-class tuned visited tables and with one table for every
    # class (slots, slots per bucket): the classes were tuned one at a time,
    # and side by side they share the CUs.  Every variant returns the same
    # results; the fastest of the side-by-side and one-kernel launches is
    # the reported one.
    from scattersearch_amd._lib import SS_CLASSES_AUTO, SS_CLASSES_OFF, SS_CLASSES_SERIAL
    variants = {}
    ref_ids = None
    best = None
    tables = [(args.visited_slots, args.visited_bucket)]
    if not args.visited_slots:
        tables += [(8192, 16), (8192, 16 | ss._lib.SS_VISITED_TWO_CHOICE), (4096, 16)]
    runs = [("classes_serial", SS_CLASSES_SERIAL, tables[0])]
    for tb in tables:
        two = tb[1] & ss._lib.SS_VISITED_TWO_CHOICE
        sfx = "" if tb == tables[0] else f"_slots{tb[0]}_bucket{tb[1] & 0xFF}" + ("_two_choice" if two else "")
        runs += [("classes_concurrent" + sfx, SS_CLASSES_AUTO, tb), ("one_kernel" + sfx, SS_CLASSES_OFF, tb)]
    for name, mode, (slots_v, bucket_v) in runs:
        wall_v, kern_v, ids, st = timed(mode, slots_v, bucket_v)
        variants[name] = round(kern_v, 4)
        log(f"mixed launch, {name}: {kern_v:.3f} ms per launch")
        if ref_ids is None:
            ref_ids, ref_hops = ids.copy(), st["hops"].copy()
        elif not np.array_equal(ids, ref_ids) or not np.array_equal(st["hops"], ref_hops):
            raise RuntimeError(f"the mixed launch ({name}) returned other results than one kernel")
        if mode != SS_CLASSES_SERIAL and (best is None or kern_v < best[2]):
            best = (name, wall_v, kern_v, ids, st, slots_v, bucket_v, mode)
    chosen, wall, kern_ms, ids, st, slots_c, bucket_c, mode_c = best
    rec = {}
    for k in (1, 10, 100):
        sel = Ks == k
        rec[f"K={k}"] = round(float(metrics.recall_per_query(ids[sel], gt[sel], k).mean()), 4)
    def launch_slots(slots):
        ix.search_device(dq.data_ptr(), nq, k_max, l_max, k_max, d_ids.data_ptr(), d_dist.data_ptr(),
                         d_stats.data_ptr(), d_K=dK.data_ptr(), d_L=dL.data_ptr(), stream=stream.cuda_stream,
                         use_graph=False, visited_slots=slots, pool_classes=SS_CLASSES_OFF)

    def stats_host():
        return d_stats.cpu().numpy().view(np.uint32).reshape(-1).view(ss.STATS_DTYPE)

    ndc_exact = exact_ndc(ids, st, launch_slots, d_ids, stats_host)
    launch_bytes = float(metrics.bytes_per_query(ndc_exact, st["hops"], w.dim, r_pad, Ks).sum())
    achieved = launch_bytes / (kern_ms * 1e-3) / 1e9
    khz = ix.wall_clock_khz()
    service_ms = (st["t_end"].astype(np.float64) - st["t_start"]) / khz
    start_ms = (st["t_start"].astype(np.float64) - float(st["t_start"].min())) / khz
    slow = int(np.argmax(service_ms))
    qps = nq * args.steps / wall
    log(f"mixed launch: {qps:.0f} QPS, {kern_ms:.3f} ms per launch, {achieved:.0f} GB/s algorithmic, recall {rec}")
    out = {
        "metric": "QPS of one launch mixing K in {1,10,100} and L uniform in K..512, 1 MI355X",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: seeded {w.centres}-component mixture (seed {w.seed}), graph built on the GPU",
        "config": {"workload": w.name, "num": ix.info()["num"], "dim": w.dim, "nq": nq, "R": w.R,
                   "metric": "L2" if w.metric == 0 else "inner_product",
                   "K_mix": {str(k): int((Ks == k).sum()) for k in (1, 10, 100)}, "L_mean": float(Ls.mean()),
                   "launch": {"pool_classes": ("concurrent" if mode_c == SS_CLASSES_AUTO
                                               else "off (one kernel, longest pool first)"),
                              "variant": chosen,
                              "visited_slots": slots_c or "tuned per class",
                              "visited_bucket": ((f"{bucket_c & 0xFF}" + (" two-choice" if bucket_c & 0x100 else ""))
                                                 if bucket_c else "tuned per class"),
                              "classes": {f"L<={b}": {k: t[k] for k in ("waves_per_block", "visited_slots",
                                                                          "visited_bucket", "team_size",
                                                                          "pool_layout")}
                                          for b, t in tuned_cls.items()}}},
        "pool_classes_kernel_ms": variants,
        "recall": rec,
        "mean_ndc": round(float(ndc_exact.mean()), 1), "mean_hops": round(float(st["hops"].mean()), 1),
        "recomputed_ndc": round(float((st["ndc"] - ndc_exact).mean()), 1),
        "latency_ms": {"service_p50": round(metrics.percentile(service_ms, 0.5), 4),
                       "service_p99": round(metrics.percentile(service_ms, 0.99), 4)},
        # the launch cannot end before its slowest search: a chain of hops
        # one after another
        "slowest_search": {"L": int(Ls[slow]), "K": int(Ks[slow]), "hops": int(st["hops"][slow]),
                           "ndc": int(ndc_exact[slow]), "service_ms": round(float(service_ms[slow]), 4),
                           "started_ms": round(float(start_ms[slow]), 4)},
        # the counters were read on one kernel longest pool first; the class
        # kernels side by side run the same searches (same distance counts
        # with the same visited table), so the record serves both
        "roofline": roofline_record(w.name, nq, "mixed", "mixed", "longest_first", achieved, launch_bytes, kern_ms,
                                    match={"visited_slots": slots_c, "visited_bucket": bucket_c}),
        "graph": {"mean_degree": ginfo["passes"][-1]["mean_degree"], "knn_s": round(ginfo["knn_s"], 1),
                  "passes": len(ginfo["passes"])},
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(base, adj, ep, queries, gt, K, L, metric, seconds):
    """ScatterSearch (MultiThreadSearchArraySimulationWithET) restated on the
    host: queries one after another, T = host threads per query; plus the
    serial search run inter-query parallel.  Bounded samples of the same
    queries, each about `seconds` of CPU wall time."""
    import oracle
    from scattersearch_amd import metrics
    T, n_aff, n_os = host_cpus()
    model = cpu_model()
    # calibrate the sample size on 20 queries
    n0 = min(20, queries.shape[0])
    t = time.perf_counter()
    oracle.scatter_search(base, adj, ep, queries[:n0], K, L, T, metric)
    per_q = (time.perf_counter() - t) / n0
    n = int(min(queries.shape[0], gt.shape[0], max(n0, seconds / max(per_q, 1e-6))))
    t = time.perf_counter()
    ids, lat, ndc = oracle.scatter_search(base, adj, ep, queries[:n], K, L, T, metric)
    wall = time.perf_counter() - t
    et_qps = n / wall
    et_rec = metrics.recall(ids, gt[:n], K)
    log(f"cpu ScatterSearch T={T}: {et_qps:.0f} QPS over {n} queries, recall {et_rec:.4f}")
    # serial search, inter-query parallel
    t = time.perf_counter()
    oracle.search_batch(base, adj, ep, queries[:n0], K, L, metric, nthreads=T)
    per_q = (time.perf_counter() - t) / n0
    n2 = int(min(queries.shape[0], gt.shape[0], max(n0, seconds / max(per_q, 1e-6))))
    t = time.perf_counter()
    sids = oracle.search_batch(base, adj, ep, queries[:n2], K, L, metric, nthreads=T)[0]
    ser_qps = n2 / (time.perf_counter() - t)
    log(f"cpu serial search x{T} threads: {ser_qps:.0f} QPS over {n2} queries")
    return {"value": round(et_qps, 1), "unit": "queries/s", "cores": T, "kind": "port",
            "sample": f"first {n} queries, L={L}, K={K}, serial, {T} threads per query (ScatterSearch)",
            "recall@10": round(et_rec, 4), "cpu": model, "cpus_allowed": n_aff, "cpus_reported": n_os,
            "serial_search_qps": round(ser_qps, 1),
            "serial_search_sample": f"first {n2} queries, {T} queries in parallel, private visited sets",
            "serial_search_recall@10": round(metrics.recall(sids, gt[:n2], K), 4)}


if __name__ == "__main__":
    main()
