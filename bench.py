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
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-entry0", action="store_true",
                    help="skip timing the FAST path from the single entry node 0 (profiling runs)")
    ap.add_argument("--seeds", type=int, default=64,
                    help="FAST-mode start nodes from k-means over the base (1 = node 0 only)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device (HBM counter calibration)")
    ap.add_argument("--workload", default="search",
                    choices=["search", "c1", "pca", "uneven", "serve", "refgraph"],
                    help="search: the headline 1M x 128 u8 batched search (default); c1: 100k x "
                         "128 u8 reference-exact build + search; pca: 1M x 960 f32 PCACompress + "
                         "search on codes; uneven: uneven batches + batched RobustPrune; serve: "
                         "10M x 96 i8 continuous-batching server under Poisson load; refgraph: "
                         "the headline search on the graph the reference's own Vamana builds")
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

    def recall_at(ids, g):
        return AT.recall_with_ties(ids, g, gtd[:len(g)], 10)

    # FAST-mode start nodes: 64 k-means seeds over the base (the plain path
    # keeps the reference's single entry, node 0)
    seeds = None
    if args.seeds > 1:
        t = time.time()
        seeds = N.pick_seeds(ix, args.seeds)
        log("%d k-means seeds in %.2fs" % (len(seeds), time.time() - t))
    tune_table = None
    if args.fast == "auto":
        t = time.time()
        best, tune_table = AT.tune(ix, q_d, gt, gtd, k=10, target=args.target, log=log,
                                   start_points=seeds)
        fast = best["config"]
        log("autotuner picked %s at L=%d in %.1fs" % (fast, best["L"], time.time() - t))
        if not args.L:
            args.L = best["L"]
    else:
        fast = tuple(int(x) for x in args.fast.split(","))
    sweep = [args.L] if args.L else [10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80, 96, 128, 160,
                                        200, 256]

    def run_fast(L, out=None):
        return N.search(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=fast, out=out,
                        start_points=seeds)

    def run_fast_entry0(L, out=None):
        return N.search(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=fast, out=out)

    def run_plain(L, out=None):
        return N.search(ix, q_d, L, 10, mode=N.SEARCH_REF, cut=1.35, out=out)

    def pick(runner, name):
        for L in sweep:
            ids, _, st = runner(L)
            torch.cuda.synchronize()
            r = recall_at(ids.cpu().numpy().view(np.uint32), gt)
            log("%s L=%d recall@10 %.4f" % (name, L, r))
            if r >= args.target:
                return L, r, st.cpu().numpy()
        return sweep[-1], r, st.cpu().numpy()

    L_fast, rec_fast, st_fast = pick(run_fast, "fast")
    sweep = [10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80, 96, 128, 160, 200, 256]
    L_plain, rec_plain, st_plain = pick(run_plain, "plain")
    cmp_e0 = seeds is not None and not args.no_entry0
    L_e0, rec_e0, _ = pick(run_fast_entry0, "fast, entry node 0") if cmp_e0 \
        else (L_fast, rec_fast, None)

    def timed(runner, L, steps, warmup, graph_of=None):
        out = runner(L)
        step = lambda: runner(L, out)  # noqa: E731
        if graph_of is not None:
            # one step captured in a HIP graph by the library itself
            # (nbpg_search_graph_*), replayed per step on torch's stream
            g = graph_of(L)
            step = g.replay
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in ev:
            a.record()
            step()
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = np.array([a.elapsed_time(b) for a, b in ev])
        return wall, per

    def graph_fast(L):
        return N.SearchGraph(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=fast, start_points=seeds)

    def graph_fast_entry0(L):
        return N.SearchGraph(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=fast)

    wall, per = timed(run_fast, L_fast, args.steps, args.warmup,
                      None if args.no_graph else graph_fast)
    nq = q.shape[0]
    qps = nq * args.steps / wall
    log("fast L=%d: %.0f QPS, step p50 %.3f ms p99 %.3f ms" %
        (L_fast, qps, np.percentile(per, 50), np.percentile(per, 99)))
    # end to end: queries from pinned host memory, search, ids and distances
    # back to pinned host memory, one synchronous batch per step (host clock)
    qh = torch.from_numpy(qpad).pin_memory()
    q_e2e = torch.empty_like(q_d)
    out_e2e = run_fast(L_fast)
    ids_h = torch.empty(tuple(out_e2e[0].shape), dtype=out_e2e[0].dtype, pin_memory=True)
    d_h = torch.empty(tuple(out_e2e[1].shape), dtype=out_e2e[1].dtype, pin_memory=True)

    def e2e_step():
        q_e2e.copy_(qh, non_blocking=True)
        N.search(ix, q_e2e, L_fast, 10, mode=N.SEARCH_FAST, fast=fast, out=out_e2e,
                 start_points=seeds)
        ids_h.copy_(out_e2e[0], non_blocking=True)
        d_h.copy_(out_e2e[1], non_blocking=True)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        e2e_step()
    e2e = []
    for _ in range(args.steps):
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
        w0, _ = timed(run_fast_entry0, L_e0, max(5, args.steps // 2), 2,
                      None if args.no_graph else graph_fast_entry0)
        e0_qps = nq * max(5, args.steps // 2) / w0
        log("fast with the single entry node 0, L=%d: %.0f QPS" % (L_e0, e0_qps))
    pwall, pper = timed(run_plain, L_plain, args.plain_steps, 1)
    plain_qps = nq * args.plain_steps / pwall
    log("plain L=%d: %.0f QPS" % (L_plain, plain_qps))

    # roofline of the fast kernel: algorithmic bytes of one launch from each
    # query's own counters -- expansions, distance computations and the
    # adjacency entries its expanded nodes hold (NBPG_SEARCH_EDGES):
    #   B_q = 4 * (visited + edges) + D * dist_cmps + D
    # The seed rows' distances (every query starts with one per seed) hit
    # the same L2-resident rows for every query; the headline figure leaves
    # them out and the figure with them is reported beside it.
    e_ids, _, st3 = N.search(ix, q_d, L_fast, 10, mode=N.SEARCH_FAST, fast=fast,
                             start_points=seeds, edges=True)
    if not np.array_equal(e_ids.cpu().numpy(), run_fast(L_fast)[0].cpu().numpy()):
        raise SystemExit("FAST results differ between two calls")
    st3 = st3.cpu().numpy().astype(np.float64)
    vis, cmps, edges = st3[:, 0], st3[:, 1], st3[:, 2]
    dbytes = base.shape[1]
    nseed = len(seeds) if seeds is not None else 1
    alg_with_seeds = float(4 * (vis + edges).sum() + (cmps * dbytes).sum() + nq * dbytes)
    alg_bytes = alg_with_seeds - float(nq * nseed * dbytes)
    kernel_ms = float(per.mean())
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    achieved_seeds = alg_with_seeds / (kernel_ms * 1e-3) / 1e9
    log("roofline: %.1f MB per launch without the seed rows (%.1f MB with), %.0f / %.0f GB/s"
        % (alg_bytes / 1e6, alg_with_seeds / 1e6, achieved, achieved_seeds))

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
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded 1000-cluster Gaussian mixture, heavy-tailed dims)",
        "config": {"workload": "1M x 128 u8 base, 10k queries, NBPG graph R=64 built on device, "
                               "k=10, fast persistent search", "n": base.shape[0], "nq": nq,
                   "L": L_fast, "fast_config": list(fast), "recall@10": round(rec_fast, 4),
                   "start_nodes": ("%d k-means seeds" % len(seeds)) if seeds is not None
                   else "node 0",
                   "hipgraph": not args.no_graph,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(base.shape[0], nq, base.shape[1], L_fast, fast),
                     "algorithmic_bytes": round(alg_bytes), "kernel_ms": round(kernel_ms, 5),
                     "bytes": "sum over queries of 4*(visited + adjacency entries of the "
                              "expanded nodes) + 128*(dist_cmps - seed rows) + 128, from the "
                              "kernel's own per-query counters",
                     "with_seed_rows": {"algorithmic_bytes": round(alg_with_seeds),
                                        "achieved": round(achieved_seeds, 1),
                                        "frac": round(achieved_seeds / PEAK_HBM_GBS, 4)}},
        "cpu_baseline": cpu,
        "p50_step_ms": round(float(np.percentile(per, 50)), 4),
        "p99_step_ms": round(float(np.percentile(per, 99)), 4),
        "end_to_end": {"qps": round(nq / (float(e2e.mean()) * 1e-3), 1),
                       "p50_ms": round(float(np.percentile(e2e, 50)), 4),
                       "p99_ms": round(float(np.percentile(e2e, 99)), 4),
                       "what": "query rows from pinned host memory, search, ids and distances "
                               "back to pinned host memory, synchronised per batch (host clock)"},
        "plain": {"L": L_plain, "qps": round(plain_qps, 1), "recall@10": round(rec_plain, 4),
                  "speedup_fast_over_plain": round(qps / plain_qps, 2)},
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
        "autotune": [{"config": list(r["config"]), "L": r["L"], "qps": round(r["qps"], 1)}
                     for r in (tune_table or [])],
    }
    print(json.dumps(res), flush=True)


GTD = None


def measured_traffic(n, nq, dims, L, fast):
    """HBM bytes per FAST launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries of this workload (calibrated against a 1 GiB copy),
    one per profiled launch shape; None when this run's workload or launch
    shape was not profiled."""
    path = os.path.join(ROOT, "profiles", "search_reg_kernel_1m_u8_L16.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except OSError:
        return None
    for p in prof["runs"]:
        if (p["n"], p["nq"], p["dims"], p["L"], list(p["fast_config"])) == \
                (n, nq, dims, L, list(fast)):
            return p["traffic_bytes_per_launch"]
    return None


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
