#!/usr/bin/env python3
"""Headline benchmark: batched beam search on 1M x 128 u8 at recall@10 >= 0.95.

One step = one persistent-kernel launch searching all query rows (10,000 by
default) with inputs already resident in HBM.  Setup (untimed): seeded
synthetic data, the NBPG graph built on the device (R=64, 2 passes, 4
splits, dynamic prune, min_alpha 1.05), exact ground truth on the device,
and a beam-width sweep that picks the smallest L reaching recall@10 >= 0.95.
Beside the fast path the plain reference-exact path (one query per
workgroup) is timed at its own smallest L with the same recall, and the
reference's own qsearchAll (compiled from its sources, oracle/_ref) is timed
on the host cores over a bounded sample as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS @ recall@10>=0.95, 1M x128 u8: batched/served vs plain GPU; p99 ms"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench %7.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


T0 = time.time()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=0, help="base rows (0 = the workload's own size)")
    ap.add_argument("--nq", type=int, default=0, help="queries (0 = the workload's own count)")
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--L", type=int, default=0, help="fixed beam width (0 = sweep)")
    ap.add_argument("--fast", default="auto",
                    help="waves_per_group,expand_width,filter_log2,groups,row_split or 'auto'")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches instead of the library's captured search graph")
    ap.add_argument("--plain-steps", type=int, default=6)
    ap.add_argument("--batches", type=int, default=12,
                    help="distinct query batches searched in turn by the timed steps")
    ap.add_argument("--streams", type=int, default=0,
                    help="streams the steps are issued on round-robin (0 = measured: the "
                         "autotuner's choice, or the best of 1..4)")
    ap.add_argument("--no-entry0", action="store_true",
                    help="skip timing the FAST path from the single entry node 0 (profiling runs)")
    ap.add_argument("--start", choices=["kmeans", "nearest"], default="kmeans",
                    help="FAST start points: kmeans = --seeds k-means seeds shared by every "
                         "query; nearest = each query's nearest of --seed-rows sampled rows "
                         "(NBPG_SEARCH_SEED_NEAREST, found by an int8-MFMA pass before the "
                         "search).  The other rule is timed beside it")
    ap.add_argument("--seed-rows", type=int, default=1024,
                    help="seed rows for the nearest-seed starts (at most 4096)")
    ap.add_argument("--seeds", type=int, default=64,
                    help="FAST-mode start nodes from k-means over the base (1 = node 0 only)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served leg (the continuous-batching server on the same index)")
    ap.add_argument("--serve-seconds", type=float, default=1.5,
                    help="seconds of request traffic per offered load in the served leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device (HBM counter calibration)")
    ap.add_argument("--workload", default="search",
                    choices=["search", "c1", "pca", "uneven", "serve", "refgraph", "traffic"],
                    help="search: the headline 1M x 128 u8 batched search (default); c1: 100k x "
                         "128 u8 reference-exact build + search; pca: 1M x 960 f32 PCACompress + "
                         "search on codes; uneven: uneven batches + batched RobustPrune; serve: "
                         "10M x 96 i8 continuous-batching server under Poisson load; refgraph: "
                         "the headline search on the graph the reference's own Vamana builds; traffic: "
                         "every FAST launch shape the autotuner can pick, for HBM counter runs")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the search would shard "
              "by queries, with the index replicated and no exchange). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "3 copies of 1 GiB", "bytes_read_per_copy": 1 << 30,
                          "bytes_written_per_copy": 1 << 30}), flush=True)
        return
    if args.workload != "search":
        import workloads  # noqa: E402  (bench-only module beside this file)
        res = getattr(workloads, "run_" + args.workload)(args, log)
        print(json.dumps(res), flush=True)
        return
    import nbpg_amd as N
    from nbpg_amd import autotune as AT
    from nbpg_amd import datagen as D

    N.init(0)
    dev = N.device_info()
    log("device", dev)
    base, q = D.workload_1m(args.n or 1_000_000, args.nq or 10_000)
    log("data", base.shape, q.shape, "zeros %.3f" % float((base == 0).mean()))
    R = 64
    ix = N.Index(N.U8, base.shape[0], base.shape[1], R)
    ix.set_points(base)
    base_d = N.to_device(base)
    stride = ix.row_stride
    qpad = np.zeros((q.shape[0], stride), np.uint8)
    qpad[:, :q.shape[1]] = q
    q_d = N.to_device(qpad)

    t = time.time()
    iso = N.build(ix, R, 70, VL=140, alpha=1.1, min_alpha=1.05, passes=2, num_split=4,
                  dynamic_prune=True, sp_seed=1)
    torch.cuda.synchronize()
    build_s = time.time() - t
    rows = ix.get_graph()
    avg_deg = float(rows[:, 0].mean())
    log("device NBPG build %.1fs, isolated %d, avg degree %.2f" % (build_s, iso, avg_deg))

    # exact top-100 ground truth: int8-MFMA scan, cross-checked by the VALU scan
    q_gt = torch.from_numpy(q).cuda()
    N.compgt(base_d, q_gt[:256], 100, N.U8, N.GT_MFMA)  # first-launch costs
    torch.cuda.synchronize()
    t = time.time()
    gt_ids, gt_d = N.compgt(base_d, q_gt, 100, N.U8, N.GT_MFMA)
    torch.cuda.synchronize()
    gt_s = time.time() - t
    t = time.time()
    p_ids, p_d = N.compgt(base_d, q_gt, 100, N.U8, N.GT_PLAIN)
    torch.cuda.synchronize()
    gt_plain_s = time.time() - t
    gt = gt_ids.cpu().numpy().view(np.uint32)
    gtd = gt_d.cpu().numpy()
    if not (np.array_equal(gt, p_ids.cpu().numpy().view(np.uint32))
            and np.array_equal(gtd, p_d.cpu().numpy())):
        raise SystemExit("MFMA and VALU ground truth differ")
    del p_ids, p_d
    # the reference's own slot-replacement rule (what its CompGT writes)
    t = time.time()
    r_ids, _ = N.compgt(base_d, q_gt, 100, N.U8, N.GT_MFMA | N.GT_REFTIES)
    torch.cuda.synchronize()
    gt_ref_s = time.time() - t
    gt_ref = r_ids.cpu().numpy().view(np.uint32)
    del r_ids
    gt_ref_differs = int((gt_ref != gt).any(axis=1).sum())
    log("device ground truth top-100: MFMA %.3fs, VALU %.3fs (identical); reference rule %.3fs,"
        " %d queries differ from the exact top-100" % (gt_s, gt_plain_s, gt_ref_s, gt_ref_differs))

    # the timed steps search B distinct query batches in turn (batch 0 is
    # the workload's own query set, the others fresh draws of the same
    # mixture), so no batch is searched twice in a row and batches that run
    # at the same time differ; recall is over all of them
    B = max(1, args.batches)
    nq = q.shape[0]
    q_host = [q] + [D.workload_1m_queries(nq, b, q.shape[1]) for b in range(1, B)]
    q_ds = [q_d]
    for qq in q_host[1:]:
        qp_ = np.zeros((nq, stride), np.uint8)
        qp_[:, :qq.shape[1]] = qq
        q_ds.append(N.to_device(qp_))
    gts, gtds = [gt], [gtd]
    for qq in q_host[1:]:
        g_i, g_d = N.compgt(base_d, torch.from_numpy(qq).cuda(), 100, N.U8, N.GT_MFMA)
        gts.append(g_i.cpu().numpy().view(np.uint32))
        gtds.append(g_d.cpu().numpy())
    log("%d query batches of %d, ground truth for each" % (B, nq))

    def recall_all(runner, L):
        rs = []
        for b in range(B):
            ids = runner(L, batch=b)[0]
            torch.cuda.synchronize()
            rs.append(AT.recall_with_ties(ids.cpu().numpy().view(np.uint32), gts[b], gtds[b], 10))
        return float(np.mean(rs))

    # FAST-mode start points (the plain path keeps the reference's single
    # entry, node 0): each query's nearest of 4096 sampled seed rows (the
    # per-query get_sp of the reference's driver, chosen on the device before
    # the search), or k-means seeds shared by every query
    seeds = None  # start points shared by every query
    fmode = N.SEARCH_FAST
    srows = np.sort(np.random.default_rng(123).choice(
        base.shape[0], min(args.seed_rows, N.MAX_SEEDS, base.shape[0]),
        replace=False)).astype(np.uint32)
    near_desc = ("each query's nearest of %d sampled seed rows (int8-MFMA pass before the "
                 "search)" % len(srows))
    # the other start rule, timed beside the headline's: (mode, shared
    # start points, description, JSON key)
    alt = None
    if args.start == "nearest":
        ix.set_seeds(srows)
        fmode |= N.SEARCH_SEED_NEAREST
        start_desc = near_desc
        if not args.no_entry0 and args.seeds > 1:
            ks = N.pick_seeds(ix, args.seeds)
            alt = (N.SEARCH_FAST, ks, "%d k-means seeds" % len(ks), "fast_kmeans_seeds")
    elif args.seeds > 1:
        t = time.time()
        seeds = N.pick_seeds(ix, args.seeds)
        log("%d k-means seeds in %.2fs" % (len(seeds), time.time() - t))
        start_desc = "%d k-means seeds" % len(seeds)
        if not args.no_entry0:
            ix.set_seeds(srows)
            alt = (N.SEARCH_FAST | N.SEARCH_SEED_NEAREST, None, near_desc,
                   "fast_nearest_seed_rows")
    else:
        start_desc = "node 0"
    log("FAST start points:", start_desc)
    tune_table = None
    streams = args.streams
    if args.fast == "auto":
        t = time.time()
        best, tune_table = AT.tune(ix, q_d, gt, gtd, k=10, target=args.target, log=log,
                                   start_points=seeds, q_batches=q_ds, mode=fmode)
        fast = best["config"]
        log("autotuner picked %s at L=%d, %d streams in %.1fs"
            % (fast, best["L"], best["streams"], time.time() - t))
        if not args.L:
            args.L = best["L"]
        if not streams:
            streams = best["streams"]
    else:
        fast = tuple(int(x) for x in args.fast.split(","))
    full_sweep = [10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80, 96, 128, 160, 200, 256]

    def run_fast(L, out=None, batch=0):
        return N.search(ix, q_ds[batch], L, 10, mode=fmode, fast=fast, out=out,
                        start_points=seeds)

    def run_fast_alt(L, out=None, batch=0):
        return N.search(ix, q_ds[batch], L, 10, mode=alt[0], fast=fast, out=out,
                        start_points=alt[1])

    def run_fast_entry0(L, out=None, batch=0):
        return N.search(ix, q_ds[batch], L, 10, mode=N.SEARCH_FAST, fast=fast, out=out)

    def run_plain(L, out=None, batch=0):
        return N.search(ix, q_ds[batch], L, 10, mode=N.SEARCH_REF, cut=1.35, out=out)

    def pick(runner, name, start=0):
        r = 0.0
        for L in [x for x in full_sweep if x >= start] or [start]:
            r = recall_all(runner, L)
            log("%s L=%d recall@10 %.4f (%d batches)" % (name, L, r, B))
            if r >= args.target:
                return L, r
        return L, r

    L_fast, rec_fast = pick(run_fast, "fast", args.L)
    L_plain, rec_plain = pick(run_plain, "plain")
    cmp_e0 = (seeds is not None or args.start == "nearest") and not args.no_entry0
    L_e0, rec_e0 = pick(run_fast_entry0, "fast, entry node 0") if cmp_e0 else (L_fast, rec_fast)
    L_alt, rec_alt = pick(run_fast_alt, "fast, " + alt[2]) if alt is not None else (None, None)

    def launchers(runner, L, S):
        """launch(i) searches batch i % B into its own output set (step i
        runs on stream i % S; steps that share a set share a stream)"""
        M = AT.output_sets(B, S)
        outs = [runner(L, batch=j % B) for j in range(M)]
        return (lambda i: runner(L, outs[i % M], batch=i % B)), outs

    def graph_launchers(L, S, starts="main"):
        # each batch's search captured by the library as a HIP graph
        # (nbpg_search_graph_*), one graph per output set; starts: the
        # headline's, node 0 ("entry0") or the other start rule ("alt")
        M = AT.output_sets(B, S)
        m = {"main": fmode, "entry0": N.SEARCH_FAST, "alt": alt[0] if alt else None}[starts]
        sp = {"main": seeds, "entry0": None, "alt": alt[1] if alt else None}[starts]
        gs = [N.SearchGraph(ix, q_ds[j % B], L, 10, mode=m, fast=fast, start_points=sp)
              for j in range(M)]
        return (lambda i: gs[i % M].replay()), gs

    def pipelined(launch, S, steps, warmup, per_batch=False):
        """steps issued round-robin over S streams, bracketed by device
        synchronisations; returns (host seconds, event-timed ms of the whole
        region, per-step ms on each step's own stream or None)"""
        ss = [torch.cuda.Stream() for _ in range(S)]
        cur = torch.cuda.current_stream()

        def issue(n, ev=None):
            for i in range(n):
                with torch.cuda.stream(ss[i % S]):
                    if ev is not None:
                        ev[i][0].record()
                    launch(i)
                    if ev is not None:
                        ev[i][1].record()
        issue(warmup)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)] if per_batch else None
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for st in ss:
            st.wait_event(e0)
        t0 = time.perf_counter()
        issue(steps, ev)
        for st in ss:
            cur.wait_stream(st)
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = np.array([a.elapsed_time(b) for a, b in ev]) if ev else None
        return wall, e0.elapsed_time(e1), per

    if not streams:
        # --fast given: pick the stream count by the same measurement
        best_s = None
        for S in (1, 2, 3, 4):
            launch, _ = launchers(run_fast, L_fast, S)
            w, _, _ = pipelined(launch, S, 24, S)
            log("fast L=%d, %d streams: %.0f QPS" % (L_fast, S, nq * 24 / w))
            if best_s is None or w < best_s[1]:
                best_s = (S, w)
        streams = best_s[0]
    S = streams
    K = args.steps
    # the timed steps: eager calls and replays of the library's graphs, each
    # over S streams; the faster is the headline (both are reported)
    launch_e, outs_e = launchers(run_fast, L_fast, S)
    wall_e, region_e, _ = pipelined(launch_e, S, K, args.warmup)
    qps_e = nq * K / wall_e
    log("fast L=%d eager, %d streams: %.0f QPS (%.4f ms per step)" % (L_fast, S, qps_e,
                                                                       wall_e / K * 1e3))
    qps_g, wall_g, region_g = None, None, None
    if not args.no_graph:
        launch_g, gs = graph_launchers(L_fast, S)
        wall_g, region_g, _ = pipelined(launch_g, S, K, args.warmup)
        qps_g = nq * K / wall_g
        log("fast L=%d graph replays, %d streams: %.0f QPS" % (L_fast, S, qps_g))
        for j in range(min(len(gs), len(outs_e))):
            if not (torch.equal(gs[j].ids, outs_e[j][0]) and torch.equal(gs[j].dists, outs_e[j][1])):
                raise SystemExit("graph replay and eager results differ (output set %d)" % j)
    use_graph = qps_g is not None and qps_g > qps_e
    wall, region, qps = (wall_g, region_g, qps_g) if use_graph else (wall_e, region_e, qps_e)
    launch_best = launch_g if use_graph else launch_e
    # per-batch latency: under the same pipelined load, and one batch at a
    # time on one stream; throughput with one stream beside it
    _, _, per = pipelined(launch_best, S, K, args.warmup, per_batch=True)
    launch_1, _ = launchers(run_fast, L_fast, 1)
    _, _, per_alone = pipelined(launch_1, 1, K, args.warmup, per_batch=True)
    wall_1, _, _ = pipelined(launch_1, 1, K, args.warmup)
    qps_1 = nq * K / wall_1
    log("fast: batch latency p50 %.3f / p99 %.3f ms with %d streams, %.3f / %.3f ms alone; "
        "one stream %.0f QPS" % (np.percentile(per, 50), np.percentile(per, 99), S,
                                 np.percentile(per_alone, 50), np.percentile(per_alone, 99), qps_1))
    # end to end: queries from pinned host memory, search, ids and distances
    # back to pinned host memory, one synchronous batch per step (host clock)
    qh = torch.from_numpy(qpad).pin_memory()
    q_e2e = torch.empty_like(q_d)
    out_e2e = run_fast(L_fast)
    ids_h = torch.empty(tuple(out_e2e[0].shape), dtype=out_e2e[0].dtype, pin_memory=True)
    d_h = torch.empty(tuple(out_e2e[1].shape), dtype=out_e2e[1].dtype, pin_memory=True)

    def e2e_step():
        q_e2e.copy_(qh, non_blocking=True)
        N.search(ix, q_e2e, L_fast, 10, mode=fmode, fast=fast, out=out_e2e,
                 start_points=seeds)
        ids_h.copy_(out_e2e[0], non_blocking=True)
        d_h.copy_(out_e2e[1], non_blocking=True)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        e2e_step()
    e2e = []
    for _ in range(K):
        t0 = time.perf_counter()
        e2e_step()
        e2e.append(time.perf_counter() - t0)
    e2e = np.array(e2e) * 1e3
    if not np.array_equal(ids_h.numpy(), out_e2e[0].cpu().numpy()):
        raise SystemExit("end-to-end results differ from the device copy")
    log("end to end (query upload, search, result download): %.0f QPS, p50 %.3f ms p99 %.3f ms"
        % (nq / (e2e.mean() * 1e-3), np.percentile(e2e, 50), np.percentile(e2e, 99)))
    e0_qps = qps
    if cmp_e0:
        K0 = max(6, K // 2)
        l0, _ = (graph_launchers(L_e0, S, starts="entry0") if use_graph
                 else launchers(run_fast_entry0, L_e0, S))
        w0, _, _ = pipelined(l0, S, K0, 2)
        e0_qps = nq * K0 / w0
        log("fast with the single entry node 0, L=%d: %.0f QPS" % (L_e0, e0_qps))
    alt_qps, alt_alone = None, None
    if alt is not None:
        la, _ = (graph_launchers(L_alt, S, starts="alt") if use_graph
                 else launchers(run_fast_alt, L_alt, S))
        wa, _, _ = pipelined(la, S, K, args.warmup)
        alt_qps = nq * K / wa
        la1, _ = launchers(run_fast_alt, L_alt, 1)
        _, _, per_a = pipelined(la1, 1, K, args.warmup, per_batch=True)
        alt_alone = float(np.percentile(per_a, 50))
        log("fast from %s, L=%d: %.0f QPS (%d streams), one batch alone p50 %.3f ms"
            % (alt[2], L_alt, alt_qps, S, alt_alone))
    lp, _ = launchers(run_plain, L_plain, S)
    pwall, _, _ = pipelined(lp, S, args.plain_steps, 1)
    plain_qps = nq * args.plain_steps / pwall
    log("plain L=%d, %d streams: %.0f QPS" % (L_plain, S, plain_qps))

    # served: the continuous-batching server (nbpg_server_*) on the same
    # index, beam width, launch shape and start points, with requests of
    # 1..64 queries (k = 10) arriving as Poisson traffic at two offered loads
    # and in a closed loop; latency is submit to completion per request,
    # host staging and result copies included
    served = None
    if not args.no_serve:
        import workloads as WL  # noqa: E402  (bench-only module beside this file)
        srv = N.Server(ix, slots=3, max_batch_queries=4096, mem_budget=256 << 20,
                       max_L=max(64, L_fast), max_k=10, fast=tuple(fast[:3]) + (0,) + tuple(fast[4:]),
                       start_points=seeds, seed_nearest=args.start == "nearest")
        loads = WL.serve_loads(srv, q, gt, gtd, L_fast,
                               [0.0025 * qps, 0.01 * qps, None, ("burst", 64)],
                               args.serve_seconds, log)
        srv.close()
        served = {"L": L_fast, "k": 10, "slots": 3, "request_queries": "1..64",
                  "start_nodes": start_desc, "loads": loads,
                  "what": "the server on this index: Poisson arrivals at 0.25% and 1% of the "
                          "batched value, a closed loop of single requests (bounded by the "
                          "Python client's submit rate), and a closed loop submitting bursts "
                          "of 64 requests per call; latency = submit to completion"}

    # roofline of the fast kernel: algorithmic bytes of one launch from each
    # query's own counters -- expansions, distance computations and the
    # adjacency entries its expanded nodes hold (NBPG_SEARCH_EDGES):
    #   B_q = 4 * (visited + edges) + D * dist_cmps + D
    # averaged over the batches.  The start rows' distances (one per shared
    # seed, or the query's nearest seed row) read rows every query's search
    # shares, from L2; the headline figure leaves them out and the figure
    # with them is reported beside it.  The nearest-seed pass before the
    # search reads its seed rows from L2 too (its time is in the step).
    dbytes = base.shape[1]
    nseed = len(seeds) if seeds is not None else 1
    st_all = []
    for b in range(B):
        e_ids, _, st3 = N.search(ix, q_ds[b], L_fast, 10, mode=fmode, fast=fast,
                                 start_points=seeds, edges=True)
        if not torch.equal(e_ids, run_fast(L_fast, batch=b)[0]):
            raise SystemExit("FAST results differ between two calls")
        st_all.append(st3.cpu().numpy().astype(np.float64))
    st3 = np.concatenate(st_all)
    vis, cmps, edges = st3[:, 0], st3[:, 1], st3[:, 2]
    alg_with_seeds = float(4 * (vis + edges).sum() + (cmps * dbytes).sum() + len(st3) * dbytes) / B
    alg_bytes = alg_with_seeds - float(nq * nseed * dbytes)
    # the timed region by HIP events (first step's stream start to the last
    # stream's end), per step: launches overlap, so this is the device time
    # each batch costs
    kernel_ms = region / K
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    achieved_seeds = alg_with_seeds / (kernel_ms * 1e-3) / 1e9
    alone_ms = float(per_alone.mean())
    log("roofline: %.1f MB per launch without the seed rows (%.1f MB with), %.0f / %.0f GB/s; "
        "one launch alone %.3f ms, %.0f GB/s" % (alg_bytes / 1e6, alg_with_seeds / 1e6, achieved,
                                                 achieved_seeds, alone_ms,
                                                 alg_bytes / (alone_ms * 1e-3) / 1e9))

    traffic, traffic_cfg = measured_traffic(base.shape[0], nq, base.shape[1], L_fast, fast,
                                            "nearest" if args.start == "nearest" else "kmeans")
    cpu = None
    if not args.no_cpu:
        global GTD
        GTD = gtd
        cpu = cpu_baseline(base, q, rows, gt, L_plain, args.cpu_seconds)

    res = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded 1000-cluster Gaussian mixture, heavy-tailed dims); %d distinct "
                "query batches searched in turn" % B,
        "config": {"workload": "1M x 128 u8 base, 10k-query batches, NBPG graph R=64 built on "
                               "device, k=10, fast persistent search", "n": base.shape[0],
                   "nq": nq, "query_batches": B, "L": L_fast, "fast_config": list(fast),
                   "recall@10": round(rec_fast, 4),
                   "start_nodes": start_desc,
                   "streams": S, "launch": "hipgraph" if use_graph else "eager",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "traffic_profiled_config": traffic_cfg,
                     "algorithmic_bytes": round(alg_bytes), "kernel_ms": round(kernel_ms, 5),
                     "kernel_ms_is": "event-timed region / steps (launches overlap on %d streams)"
                                     % S,
                     "bytes": "mean over the batches of the sum over queries of 4*(visited + "
                              "adjacency entries of the expanded nodes) + 128*(dist_cmps - seed "
                              "rows) + 128, from the kernel's own per-query counters",
                     "with_seed_rows": {"algorithmic_bytes": round(alg_with_seeds),
                                        "achieved": round(achieved_seeds, 1),
                                        "frac": round(achieved_seeds / PEAK_HBM_GBS, 4)},
                     "one_launch_alone": {"kernel_ms": round(alone_ms, 5),
                                          "achieved": round(alg_bytes / (alone_ms * 1e-3) / 1e9,
                                                            1),
                                          "frac": round(alg_bytes / (alone_ms * 1e-3) / 1e9
                                                        / PEAK_HBM_GBS, 4)}},
        "cpu_baseline": cpu,
        "p50_step_ms": round(float(np.percentile(per, 50)), 4),
        "p99_step_ms": round(float(np.percentile(per, 99)), 4),
        "step_ms_is": "one batch's device time on its stream, %d streams in flight" % S,
        "alone": {"p50_ms": round(float(np.percentile(per_alone, 50)), 4),
                  "p99_ms": round(float(np.percentile(per_alone, 99)), 4),
                  "qps_one_stream": round(qps_1, 1)},
        "pipelined": {"eager_qps": round(qps_e, 1),
                      "graph_qps": round(qps_g, 1) if qps_g is not None else None},
        "end_to_end": {"qps": round(nq / (float(e2e.mean()) * 1e-3), 1),
                       "p50_ms": round(float(np.percentile(e2e, 50)), 4),
                       "p99_ms": round(float(np.percentile(e2e, 99)), 4),
                       "what": "query rows from pinned host memory, search, ids and distances "
                               "back to pinned host memory, synchronised per batch (host clock)"},
        "served": served,
        "plain": {"L": L_plain, "qps": round(plain_qps, 1), "recall@10": round(rec_plain, 4),
                  "streams": S, "speedup_fast_over_plain": round(qps / plain_qps, 2)},
        (alt[3] if alt else "fast_other_start"): {
            "start_nodes": alt[2], "L": L_alt, "qps": round(alt_qps, 1),
            "recall@10": round(rec_alt, 4), "alone_p50_ms": round(alt_alone, 4)}
        if alt_qps is not None else None,
        "fast_entry_node0": {"L": L_e0, "qps": round(e0_qps, 1), "recall@10": round(rec_e0, 4)}
        if cmp_e0 else None,
        "build_s": round(build_s, 2),
        "groundtruth_s": round(gt_s, 3),
        "groundtruth_valu_s": round(gt_plain_s, 3),
        "groundtruth_refrule_s": round(gt_ref_s, 3),
        "groundtruth_refrule_queries_differing": gt_ref_differs,
        "mean_visited": round(float(vis.mean()), 1),
        "mean_dist_cmps": round(float(cmps.mean()), 1),
        "mean_expanded_degree": round(float(edges.sum() / max(vis.sum(), 1.0)), 2),
        "graph_avg_degree": round(avg_deg, 2),
        "autotune": [{"config": list(r["config"]), "L": r["L"], "streams": r["streams"],
                      "qps": round(r["qps"], 1)} for r in (tune_table or [])],
    }
    print(json.dumps(res), flush=True)


GTD = None


def measured_traffic(n, nq, dims, L, fast, start):
    """HBM bytes per FAST launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries of this workload (calibrated against a 1 GiB copy),
    one per profiled launch shape: (bytes, profiled shape).  A shape that
    differs only in its group count takes the bytes of the profiled one (the
    grid size changes occupancy, not what a query reads: 2048 and 3072
    groups fetched within 0.3% of each other).  start: "nearest" (each
    query's nearest seed row) or "kmeans" (64 shared k-means seeds); the
    profiled run must have used the same.  (None, None) when this workload,
    start rule or kernel shape was not profiled."""
    path = os.path.join(ROOT, "profiles", "search_reg_kernel_1m_u8_L16.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except OSError:
        return None, None
    same = [p for p in prof["runs"] if (p["n"], p["nq"], p["dims"], p["L"]) == (n, nq, dims, L)
            and p.get("start", "").startswith(start if start == "nearest" else "64 k-means")]
    for p in same:
        if list(p["fast_config"]) == list(fast):
            return p["traffic_bytes_per_launch"], list(fast)
    for p in same:
        c = list(p["fast_config"])
        if c[:3] + c[4:] == list(fast)[:3] + list(fast)[4:]:
            return p["traffic_bytes_per_launch"], c
    return None, None


def cpu_baseline(base, q, rows, gt, L, seconds):
    """The reference's qsearchAll on the host cores, bounded sample."""
    # threads pinned to cores, one per core: read by the host OpenMP runtime
    # when the reference library loads (torch carries its own runtime)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    try:
        import oracle as O
    except Exception as e:  # pragma: no cover
        log("cpu baseline unavailable:", e)
        return None
    # the reference's build flags (-Ofast -funroll-loops ..., at the host's
    # ISA level); the parity build when the timing build is missing
    R, isa = O.ref_fast()
    flags = "-Ofast -fopenmp -fopenmp-simd -funroll-loops -falign-jumps=64 -mcx16 " \
            "-march=x86-64-%s" % isa
    if R is None:
        R = O.ref()
        flags = "-O3 -fopenmp -ffp-contract=off (parity build)"
    kind = "reference"
    if R is None:
        kind = "port"
        flags = "oracle -O2 -fopenmp"
    k = 10
    if kind == "reference":
        h = R.nbpgr_open(base.ctypes.data, base.shape[0], base.shape[1], rows.ctypes.data,
                         rows.shape[1] - 1)
        threads = R.nbpgr_threads()

        def run(nq):
            qq = np.ascontiguousarray(q[:nq])
            ids = np.zeros((nq, k), np.uint32)
            st = np.zeros((nq, 2), np.uint32)
            s = R.nbpgr_timed_search(h, qq.ctypes.data, nq, L, k, 1.35, base.shape[0],
                                     rows.shape[1] - 1, 0, ids.ctypes.data, st.ctypes.data)
            return s, ids
    else:
        threads = os.cpu_count()

        def run(nq):
            t = time.perf_counter()
            ids, _, _ = O.search(rows, base, q[:nq], O.U8, L, k)
            return time.perf_counter() - t, ids
    # whole query batches, repeated until about `seconds` of host time
    nq = len(q)
    total, reps, ids = 0.0, 0, None
    while total < seconds or reps == 0:
        s, ids = run(nq)
        total += s
        reps += 1
    if kind == "reference":
        R.nbpgr_close(h)
    from nbpg_amd.autotune import recall_with_ties
    rec = recall_with_ties(ids, gt[:nq], GTD[:nq], 10)
    log("cpu %s: %d x %d queries in %.2fs on %d threads, recall %.4f" %
        (kind, reps, nq, total, threads, rec))
    model, _ = O.host_cpu()
    return {"value": round(reps * nq / total, 1), "unit": "queries/s", "cores": int(threads),
            "kind": kind, "recall@10": round(rec, 4), "flags": flags, "cpu": model,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES"),
            "sample": "%d passes over the %d queries (%.1f s), qsearchAll at L=%d, k=10, cut 1.35, "
                      "same graph" % (reps, nq, total, L)}


if __name__ == "__main__":
    main()
