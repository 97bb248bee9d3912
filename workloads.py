"""The non-headline workloads of bench.py (--workload c1 | pca | uneven | serve | refgraph | traffic).

Each run_* function sets up its workload on the device, measures it, checks
it against the reference or the oracle where that is cheap, and returns one
JSON-able dict.  Reference binaries and the extern "C" shim come from
oracle/_ref (compiled from the reference's sources in the dev container);
the oracle is used only as a checker and CPU baseline here.
"""
import os
import subprocess
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0


def _torch():
    import torch
    return torch


def events_ms(fn, reps, warm=1):
    """Mean device time of fn() over reps calls (events on the current stream)."""
    torch = _torch()
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def replay_ms(graph, reps):
    """Mean device time of a library search graph's replay (nbpg_search_graph_*
    or the uneven form) over reps replays; the graph is closed afterwards."""
    try:
        return events_ms(graph.replay, reps)
    finally:
        graph.close()


def pad_rows(a, stride):
    out = np.zeros((a.shape[0], stride), np.uint8)
    out[:, :a.shape[1]] = a.view(np.uint8)
    return out


def recall(ids, gt, gtd, k):
    from nbpg_amd.autotune import recall_with_ties
    return recall_with_ties(np.asarray(ids)[:, :k], gt, gtd, k)


def ref_lib():
    try:
        import oracle as O
        return O.ref()
    except Exception:
        return None


def ref_timing_lib():
    """The reference's code for CPU timing: built with its own optimisation
    flags at this host's ISA level when that build exists, else the parity
    build.  Returns (library, flags) or (None, None)."""
    try:
        import oracle as O
    except Exception:
        return None, None
    R, isa = O.ref_fast()
    if R is not None:
        return R, ("-Ofast -fopenmp -fopenmp-simd -funroll-loops -falign-jumps=64 -mcx16 "
                   "-march=x86-64-%s" % isa)
    R = O.ref()
    return R, ("-O3 -fopenmp -ffp-contract=off (parity build)" if R is not None else None)


def write_bin(path, a):
    with open(path, "wb") as f:
        np.array([a.shape[0], a.shape[1]], np.uint32).tofile(f)
        np.ascontiguousarray(a).tofile(f)


def read_graph(path):
    raw = np.fromfile(path, dtype=np.uint32)
    n, R = int(raw[0]), int(raw[1])
    deg = raw[2:2 + n]
    rows = np.zeros((n, R + 1), np.uint32)
    rows[:, 0] = deg
    nb = raw[2 + n:]
    starts = np.concatenate([[0], np.cumsum(deg.astype(np.int64))])
    col = np.arange(R)[None, :]
    mask = col < deg[:, None]
    idx = (starts[:-1, None] + col)[mask]
    rows[:, 1:][mask] = nb[idx]
    return rows


def live(rows):
    out = rows.copy()
    out[:, 1:][np.arange(rows.shape[1] - 1)[None, :] >= rows[:, :1]] = 0
    return out


# ---------------------------------------------------------------------- c1
def run_c1(args, log):
    """100k x 128 u8, R32 L75 alpha 1.2 one pass: device build vs the
    reference's Vamana binary row for row, plain (reference-exact) search
    IDs vs the reference's qsearchAll, and fast search QPS / recall."""
    torch = _torch()
    import nbpg_amd as N
    from nbpg_amd import datagen as D
    N.init(0)
    n, nq = args.n or 100_000, args.nq or 1_000
    base, q = D.workload_small(n, nq)
    ix = N.Index(N.U8, n, base.shape[1], 32)
    ix.set_points(base)
    torch.cuda.synchronize()
    t = time.time()
    N.build(ix, 32, 75, alpha=1.2, passes=1, num_split=0)
    torch.cuda.synchronize()
    build_s = time.time() - t
    rows = ix.get_graph()
    log("c1 device build %.2fs" % build_s)
    res = {"workload": "c1", "n": n, "nq": nq, "device_build_s": round(build_s, 3)}
    exe = os.path.join(ROOT, "oracle", "_ref", "Vamana")
    if os.path.exists(exe) and not args.no_cpu:
        tmp = tempfile.mkdtemp()
        write_bin(os.path.join(tmp, "b.u8bin"), base)
        t = time.time()
        subprocess.run([exe, "--quant_base_path", os.path.join(tmp, "b.u8bin"), "--index_path",
                        os.path.join(tmp, "g"), "--rebuild", "--R", "32", "--L", "75",
                        "--max_alpha", "1.2", "--num_split", "0", "--num_passes", "1"],
                       check=True, capture_output=True)
        ref_s = time.time() - t
        ref_rows = read_graph(os.path.join(tmp, "g.graph"))
        same = bool((live(ref_rows) == live(rows)).all())
        rows_diff = int((live(ref_rows) != live(rows)).any(axis=1).sum())
        log("c1 reference Vamana build %.1fs on the host; graphs identical: %s (%d rows differ)"
            % (ref_s, same, rows_diff))
        res.update({"reference_build_s": round(ref_s, 2), "graph_identical": same,
                    "graph_rows_differ": rows_diff, "build_speedup": round(ref_s / build_s, 1)})
    base_d = N.to_device(base)
    q_t = torch.from_numpy(q).cuda()
    gt_ids, gt_d = N.compgt(base_d, q_t, 100, N.U8, N.GT_MFMA)
    gt, gtd = gt_ids.cpu().numpy().view(np.uint32), gt_d.cpu().numpy()
    q_d = N.to_device(pad_rows(q, ix.row_stride))
    R = ref_lib()
    h = R.nbpgr_open(base.ctypes.data, n, base.shape[1], rows.ctypes.data, 32) if R else None
    per_L = []
    for L in (16, 32, 64, 128):
        pi, _, pst = N.search(ix, q_d, L, 10, mode=N.SEARCH_REF)
        pi = pi.cpu().numpy().view(np.uint32)
        plain_ms = events_ms(lambda: N.search(ix, q_d, L, 10, mode=N.SEARCH_REF), 5)
        out = N.search(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=(2, 1, 11, 0, 0))
        fast_ms = replay_ms(N.SearchGraph(ix, q_d, L, 10, mode=N.SEARCH_FAST,
                                          fast=(2, 1, 11, 0, 0)), 20)
        fi = out[0].cpu().numpy().view(np.uint32)
        row = {"L": L, "plain_qps": round(nq / plain_ms * 1e3, 1),
               "fast_qps": round(nq / fast_ms * 1e3, 1),
               "plain_recall@10": round(recall(pi, gt, gtd, 10), 4),
               "fast_recall@10": round(recall(fi, gt, gtd, 10), 4)}
        if h:
            ids = np.zeros((nq, 10), np.uint32)
            st = np.zeros((nq, 2), np.uint32)
            secs = R.nbpgr_timed_search(h, q.ctypes.data, nq, L, 10, 1.35, n, 32, 0,
                                        ids.ctypes.data, st.ctypes.data)
            row["reference_id_match"] = round(float((ids == pi).all(axis=1).mean()), 4)
            row["reference_stats_match"] = bool((st == pst.cpu().numpy().view(np.uint32)).all())
            row["cpu_reference_qps"] = round(nq / secs, 1)
        log("c1", row)
        per_L.append(row)
    if h:
        R.nbpgr_close(h)
        res["cpu_threads"] = int(R.nbpgr_threads())
    res["per_L"] = per_L
    ix.close()
    return res


# ---------------------------------------------------------------------- pca
def gt_float(x, xq, k, chunk=1 << 17):
    """Exact fp32 L2 top-k of each xq row over x (device, chunked)."""
    torch = _torch()
    xn = (x * x).sum(1)
    best_d = torch.full((xq.shape[0], k), float("inf"), device=x.device)
    best_i = torch.zeros((xq.shape[0], k), dtype=torch.int64, device=x.device)
    for s in range(0, x.shape[0], chunk):
        e = min(x.shape[0], s + chunk)
        d = xn[s:e][None, :] - 2.0 * (xq @ x[s:e].T)
        dd, ii = torch.topk(d, min(k, e - s), dim=1, largest=False)
        cd = torch.cat([best_d, dd], 1)
        ci = torch.cat([best_i, ii + s], 1)
        o = torch.argsort(cd, dim=1)[:, :k]
        best_d, best_i = torch.gather(cd, 1, o), torch.gather(ci, 1, o)
    return best_i.cpu().numpy().astype(np.uint32), best_d


def run_pca(args, log):
    """1M x 960 f32 -> PCA 64: FP64 statistics, host eigendecomposition,
    projection (plain fp32 and bf16 MFMA) with its error against FP64, codes
    (i8, u4x2, i4x2), a Vamana build (R48, alpha 1.2), fast search on the
    codes and the same search re-scored with the fp32 rows (rerank mode)."""
    torch = _torch()
    import nbpg_amd as N
    from nbpg_amd import datagen as D
    from nbpg_amd import pca as P
    N.init(0)
    n, nq, d_in, d_out = args.n or 1_000_000, args.nq or 10_000, 960, 64
    x, xq = D.workload_pca(n, nq, d_in)
    torch.cuda.synchronize()
    res = {"workload": "pca", "n": n, "nq": nq, "d_in": d_in, "d_out": d_out}
    stats_ms = events_ms(lambda: N.pca_stats(x), 2)
    t = time.time()
    f = P.fit(x, d_out)
    res["fit_s"] = round(time.time() - t, 3)
    res["stats_f64_ms"] = round(stats_ms, 3)
    ev = f["eigenvalues"]
    res["variance_kept"] = round(float(ev[:d_out].sum() / ev.sum()), 4)
    W = torch.from_numpy(f["W"]).cuda()
    b = torch.from_numpy(f["bias"]).cuda()
    log("pca stats %.2f ms, fit %.2fs, variance kept %.3f" % (stats_ms, res["fit_s"],
                                                               res["variance_kept"]))
    proj = {}
    for mode, mname in ((N.PCA_PLAIN, "plain_fp32"), (N.PCA_BF16, "bf16_mfma")):
        for code, cname, chunk in ((N.CODE_I8, "i8", 1), (N.CODE_U4, "u4x2", 2)):
            ms = events_ms(lambda: N.pca_project(x, W, b, code=code, mode=mode), 3)
            nbytes = 4.0 * n * d_in + n * d_out / chunk + 4.0 * d_out * d_in
            gbs = nbytes / (ms * 1e-3) / 1e9
            proj["%s_%s" % (mname, cname)] = {"ms": round(ms, 3), "GBs": round(gbs, 1),
                                              "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}
            log("pca project %s -> %s: %.3f ms, %.0f GB/s" % (mname, cname, ms, gbs))
    res["projection"] = proj
    # error against an FP64 projection on a 200k-row sample
    S = min(n, 200_000)
    xs = x[:S]
    y64 = xs.double() @ W.double().T + b.double()
    _, yb = N.pca_project(xs, W, b, code=N.CODE_F32, mode=N.PCA_BF16)
    _, yp = N.pca_project(xs, W, b, code=N.CODE_F32, mode=N.PCA_PLAIN)
    eb = (yb.double() - y64).abs()
    ep = (yp.double() - y64).abs()
    res["bf16_err_vs_fp64"] = {"max": float(eb.max()), "mean": float(eb.mean())}
    res["fp32_err_vs_fp64"] = {"max": float(ep.max()), "mean": float(ep.mean())}
    c64 = N.pca_quantize(y64.float(), N.CODE_I8)
    cb, _ = N.pca_project(xs, W, b, code=N.CODE_I8, mode=N.PCA_BF16)
    res["bf16_i8_codes_differing_frac"] = float((cb != c64).float().mean())
    c64 = N.pca_quantize(y64.float(), N.CODE_U4)
    cb, _ = N.pca_project(xs, W, b, code=N.CODE_U4, mode=N.PCA_BF16)
    res["bf16_u4_bytes_differing_frac"] = float((cb != c64).float().mean())
    log("pca bf16 err max %.3g mean %.3g" % (res["bf16_err_vs_fp64"]["max"],
                                              res["bf16_err_vs_fp64"]["mean"]))
    del y64, yb, yp, eb, ep
    # ground truth in the original space
    gt_f, _ = gt_float(x, xq, 100)
    # search on the codes.  The scale of each fit maps a quantile of the
    # sample's |projection| to 1 (pca.fit's clip): 0.9999 for i8 and 0.995
    # for u4x2 codes (16 steps), chosen by the exact code-space top-10
    # against the float ground truth on a 200k-row sample.  i4x2 keeps the
    # largest value: its distance reads signed nibbles as unsigned
    # (point.h:107-120), so -1 and 0 are 15 steps apart and saturating more
    # values near zero only adds such jumps
    searches = []
    clips = {"i8": 0.9999, "u4x2": 0.995, "i4x2": 1.0}
    for code, dt, name in ((N.CODE_I8, N.I8, "i8"), (N.CODE_U4, N.U4X2, "u4x2"),
                           (N.CODE_I4, N.I4X2, "i4x2")):
        fc = P.fit(x, d_out, clip=clips[name])
        Wc, bc = torch.from_numpy(fc["W"]).cuda(), torch.from_numpy(fc["bias"]).cuda()
        cx, _ = N.pca_project(x, Wc, bc, code=code, mode=N.PCA_BF16)
        cq, _ = N.pca_project(xq, Wc, bc, code=code, mode=N.PCA_BF16)
        cx_h = cx.cpu().numpy()
        ix = N.Index(dt, n, cx_h.shape[1], 48)
        ix.set_points(cx_h)
        torch.cuda.synchronize()
        t = time.time()
        N.build(ix, 48, 96, alpha=1.2, passes=1, num_split=0)
        torch.cuda.synchronize()
        bs = time.time() - t
        cgt, cgd = N.compgt(cx, cq, 100, dt, N.GT_MFMA if dt == N.I8 else N.GT_PLAIN)
        cgt, cgd = cgt.cpu().numpy().view(np.uint32), cgd.cpu().numpy()
        q_d = N.to_device(pad_rows(cq.cpu().numpy(), ix.row_stride))
        rows = []
        for L in (16, 32, 64, 128):
            out = N.search(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=(2, 1, 11, 0, 0))
            ms = replay_ms(N.SearchGraph(ix, q_d, L, 10, mode=N.SEARCH_FAST,
                                         fast=(2, 1, 11, 0, 0)), 10)
            ids = out[0].cpu().numpy().view(np.uint32)
            rows.append({"L": L, "qps": round(nq / ms * 1e3, 1),
                         "recall@10_code_space": round(recall(ids, cgt, cgd, 10), 4),
                         "recall@10_vs_float_gt": round(float(
                             np.mean([len(set(ids[i]) & set(gt_f[i, :10])) / 10.0
                                      for i in range(nq)])), 4)})
            # rerank mode: the beam re-scored with the fp32 rows (all L of
            # it at the reference's default rerank_factor 100)
            rr = lambda: N.search_rerank(ix, q_d, x, xq, L, 10, 100, mode=N.SEARCH_FAST,
                                         fast=(2, 1, 11, 0, 0))
            rids = rr()[0].cpu().numpy().view(np.uint32)
            ms_r = events_ms(rr, 5)
            rows[-1].update({"qps_rerank_f32": round(nq / ms_r * 1e3, 1),
                             "recall@10_rerank_vs_float_gt": round(float(
                                 np.mean([len(set(rids[i]) & set(gt_f[i, :10])) / 10.0
                                          for i in range(nq)])), 4)})
        exact_code = round(float(np.mean([len(set(cgt[i, :10]) & set(gt_f[i, :10])) / 10.0
                                          for i in range(nq)])), 4)
        searches.append({"codes": name, "scale_clip": clips[name], "build_s": round(bs, 2),
                         "exact_code_top10_vs_float_gt": exact_code, "per_L": rows})
        log("pca search on %s codes: build %.1fs, %s" % (name, bs, rows))
        ix.close()
    res["search_on_codes"] = searches
    # CPU baseline: the reference's matmul + compress on a bounded sample
    R, cpu_flags = ref_timing_lib()
    if R is not None and not args.no_cpu:
        m = 20_000
        xh = x[:m].cpu().numpy()
        Wh, bh = f["W"], f["bias"]
        y = np.zeros((m, d_out), np.float32)
        t = time.perf_counter()
        R.nbpgr_matmul(xh.ctypes.data, m, d_in, Wh.ctypes.data, bh.ctypes.data, d_out,
                       y.ctypes.data)
        out = np.zeros((m, d_out), np.uint8)
        R.nbpgr_compress(y.ctypes.data, m, d_out, 1, 1, out.ctypes.data)
        cpu_s = time.perf_counter() - t
        res["cpu_baseline"] = {"rows_per_s": round(m / cpu_s, 1), "kind": "reference",
                               "cores": int(R.nbpgr_threads()), "flags": cpu_flags,
                               "sample": "%d rows, matmul + compress to i8" % m}
        # bit parity against the parity build (the timing build may reorder
        # its float sums)
        Rp = ref_lib()
        if Rp is not None:
            y = np.zeros((m, d_out), np.float32)
            Rp.nbpgr_matmul(xh.ctypes.data, m, d_in, Wh.ctypes.data, bh.ctypes.data, d_out,
                            y.ctypes.data)
            _, yp = N.pca_project(x[:m], W, b, code=N.CODE_F32, mode=N.PCA_PLAIN)
            res["plain_fp32_equals_reference_y"] = bool((yp.cpu().numpy() == y).all())
    return res


# ---------------------------------------------------------------------- traffic
def run_traffic(args, log):
    """HBM traffic per launch of every FAST launch shape the headline's
    autotuner can pick, for a rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE run:
    the headline's data, device-built graph, start points (--start) and 12
    query batches; per shape (waves per group, expansion width, filter size,
    row split; groups = as many as the device holds) the smallest beam width
    with recall@10 >= 0.95 on batch 0, then, after every shape's beam width
    is chosen, one eager search of each batch per shape in the order listed
    in "launches" -- the last len(launches) x batches search dispatches of
    the run, one per (shape, batch)."""
    torch = _torch()
    import itertools
    import nbpg_amd as N
    from nbpg_amd import autotune as AT
    from nbpg_amd import datagen as D
    N.init(0)
    n, nq = args.n or 1_000_000, args.nq or 10_000
    base, q = D.workload_1m(n, nq)
    ix = N.Index(N.U8, n, base.shape[1], 64)
    ix.set_points(base)
    N.build(ix, 64, 70, VL=140, alpha=1.1, min_alpha=1.05, passes=2, num_split=4,
            dynamic_prune=True, sp_seed=1)
    base_d = N.to_device(base)
    B = max(1, args.batches)
    q_ds = [N.to_device(pad_rows(q, ix.row_stride))]
    for b in range(1, B):
        q_ds.append(N.to_device(pad_rows(D.workload_1m_queries(nq, b, q.shape[1]),
                                         ix.row_stride)))
    gt_ids, gt_d = N.compgt(base_d, torch.from_numpy(q).cuda(), 100, N.U8, N.GT_MFMA)
    gt, gtd = gt_ids.cpu().numpy().view(np.uint32), gt_d.cpu().numpy()
    # the headline's start points (bench.py --start): each query's nearest of
    # 4096 sampled seed rows, or 64 k-means seeds shared by every query
    seeds, fmode = None, N.SEARCH_FAST
    if args.start == "nearest":
        ix.set_seeds(np.sort(np.random.default_rng(123).choice(
            n, min(args.seed_rows, N.MAX_SEEDS, n), replace=False)).astype(np.uint32))
        fmode |= N.SEARCH_SEED_NEAREST
        start = "nearest of %d seed rows" % min(args.seed_rows, N.MAX_SEEDS, n)
    else:
        seeds = N.pick_seeds(ix, args.seeds)
        start = "%d k-means seeds" % len(seeds)
    sp = AT.DEFAULT_SPACE
    shapes = []
    for wpg, w, fl, rs in itertools.product(sp["waves_per_group"], sp["expand_width"],
                                            sp["filter_log2"], sp["row_split"]):
        cfg = (wpg, w, fl, 0, rs)
        for L in AT.DEFAULT_L:
            if L < 10:
                continue
            ids = N.search(ix, q_ds[0], L, 10, mode=fmode, fast=cfg, stats=False,
                           start_points=seeds)[0]
            r = AT.recall_with_ties(ids.cpu().numpy().view(np.uint32), gt, gtd, 10)
            if r >= args.target:
                shapes.append((cfg, L))
                break
    torch.cuda.synchronize()
    log("traffic: %d shapes, %d launches each" % (len(shapes), B))
    launches = []
    for cfg, L in shapes:
        for b in range(B):
            N.search(ix, q_ds[b], L, 10, mode=fmode, fast=cfg, start_points=seeds)
        torch.cuda.synchronize()
        launches.append({"fast_config": list(cfg), "L": L})
    ix.close()
    return {"workload": "traffic", "n": n, "nq": nq, "dims": int(base.shape[1]),
            "query_batches": B, "start": start, "launches": launches}


# ---------------------------------------------------------------------- refgraph
def run_refgraph(args, log):
    """The headline search (1M x 128 u8, 10k queries, k = 10) on the NBPG
    graph built by the reference's own Vamana binary on the host cores
    (README.md:96 flags: --R 64 --num_split 4 --num_passes 2
    --dynamic_prune --min_alpha 1.05; L 70, VL 140, max_alpha 1.1 by
    default, Vamana.cpp:24-50), compiled with its own optimisation flags,
    written as a .graph file and loaded by nbpg_index_load_graph; beside it
    the device-built graph of the same flags.  FAST and plain QPS at the
    smallest L reaching recall@10 >= 0.95 on each graph."""
    torch = _torch()
    import oracle as O
    import nbpg_amd as N
    from nbpg_amd import datagen as D
    N.init(0)
    n, nq = args.n or 1_000_000, args.nq or 10_000
    base, q = D.workload_1m(n, nq)
    _, isa = O.host_cpu()
    exe = os.path.join(ROOT, "oracle", "_ref", "Vamana_fast_%s" % isa) if isa else None
    if exe is None or not os.path.exists(exe):
        exe = os.path.join(ROOT, "oracle", "_ref", "Vamana")
    if not os.path.exists(exe):
        raise SystemExit("the reference's Vamana binary is not built (oracle/Makefile ref)")
    tmp = tempfile.mkdtemp()
    write_bin(os.path.join(tmp, "b.u8bin"), base)
    t = time.time()
    proc = subprocess.Popen([exe, "--quant_base_path", os.path.join(tmp, "b.u8bin"),
                             "--index_path", os.path.join(tmp, "ref"), "--R", "64",
                             "--num_split", "4", "--num_passes", "2", "--dynamic_prune",
                             "--min_alpha", "1.05", "--rebuild"],
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    for line in proc.stdout:  # the binary's own progress lines keep the run visible
        log("ref Vamana:", line.rstrip()[:160])
    if proc.wait() != 0:
        raise SystemExit("reference Vamana failed")
    ref_build_s = time.time() - t
    log("reference build %.1fs (%s)" % (ref_build_s, os.path.basename(exe)))
    bd = N.to_device(base)
    gi, gd = N.compgt(bd, torch.from_numpy(q).cuda(), 100, N.U8, N.GT_MFMA)
    gt, gtd = gi.cpu().numpy().view(np.uint32), gd.cpu().numpy()
    qd = N.to_device(pad_rows(q, 128))
    # timing batches: the query set and three fresh draws of the same mixture
    qds = [qd] + [N.to_device(pad_rows(D.workload_1m_queries(nq, b), 128)) for b in (1, 2, 3)]
    from nbpg_amd import autotune as AT
    res = {"workload": "refgraph", "n": n, "nq": nq, "reference_build_s": round(ref_build_s, 1),
           "reference_binary": os.path.basename(exe), "graphs": []}
    for name in ("reference", "device"):
        ix = N.Index(N.U8, n, 128, 64)
        ix.set_points(base)
        if name == "reference":
            ix.load_graph(os.path.join(tmp, "ref.graph"))
            build_s = ref_build_s
        else:
            t = time.time()
            N.build(ix, 64, 70, VL=140, alpha=1.1, min_alpha=1.05, passes=2, num_split=4,
                    dynamic_prune=True, sp_seed=1)
            torch.cuda.synchronize()
            build_s = time.time() - t
        rows = ix.get_graph()
        seeds = N.pick_seeds(ix, 64)
        row = {"graph": name, "build_s": round(build_s, 2),
               "avg_degree": round(float(rows[:, 0].mean()), 2)}
        for mode, tag, sp in ((N.SEARCH_FAST, "fast", seeds), (N.SEARCH_REF, "plain", None)):
            for L in (10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80, 96, 128):
                out = N.search(ix, qd, L, 10, mode=mode, start_points=sp,
                               fast=(2, 1, 10, 0, 0))
                r = recall(out[0].cpu().numpy().view(np.uint32), gt, gtd, 10)
                if r >= 0.95:
                    break
            reps = 10 if mode == N.SEARCH_FAST else 3
            ms = events_ms(lambda: N.search(ix, qd, L, 10, mode=mode, start_points=sp,
                                            fast=(2, 1, 10, 0, 0), out=out), reps)
            # consecutive batches (the four query sets in turn) on 4 streams
            M = AT.output_sets(len(qds), 4)
            outs = [N.search(ix, qds[j % 4], L, 10, mode=mode, start_points=sp,
                             fast=(2, 1, 10, 0, 0)) for j in range(M)]
            ms4 = AT.pipelined_ms(lambda i: N.search(ix, qds[i % 4], L, 10, mode=mode,
                                                     start_points=sp, fast=(2, 1, 10, 0, 0),
                                                     out=outs[i % M]), 4, 24 if reps == 10 else 8)
            row[tag] = {"L": L, "recall@10": round(r, 4), "qps": round(nq / ms * 1e3, 1),
                        "qps_4_streams": round(nq / ms4 * 1e3, 1)}
        row["fast_over_plain"] = round(row["fast"]["qps"] / row["plain"]["qps"], 2)
        row["fast_over_plain_4_streams"] = round(row["fast"]["qps_4_streams"]
                                                 / row["plain"]["qps_4_streams"], 2)
        log("refgraph", row)
        res["graphs"].append(row)
        ix.close()
    return res


# ---------------------------------------------------------------------- uneven
def run_uneven(args, log):
    """1M x 128 u8 NBPG graph: batches of 1..4096 queries with per-query L
    log-uniform in [16, 256] and k in {1, 10, 100}, one hipGraph-captured
    launch each, autotuned launch shape; batched RobustPrune over 100k nodes
    x 256 candidates checked against the reference's prune."""
    torch = _torch()
    import nbpg_amd as N
    from nbpg_amd import datagen as D
    N.init(0)
    n, nq = args.n or 1_000_000, args.nq or 10_000
    base, q = D.workload_1m(n, nq)
    ix = N.Index(N.U8, n, base.shape[1], 64)
    ix.set_points(base)
    N.build(ix, 64, 70, VL=140, alpha=1.1, min_alpha=1.05, passes=2, num_split=4,
            dynamic_prune=True, sp_seed=1)
    base_d = N.to_device(base)
    gt_ids, gt_d = N.compgt(base_d, torch.from_numpy(q).cuda(), 100, N.U8, N.GT_MFMA)
    gt, gtd = gt_ids.cpu().numpy().view(np.uint32), gt_d.cpu().numpy()
    qpad = pad_rows(q, ix.row_stride)
    rng = np.random.default_rng(17)
    Lq = np.exp(rng.uniform(np.log(16), np.log(256), nq)).round().astype(np.int32)
    kq = rng.choice(np.array([1, 10, 100], np.int32), nq)
    kq = np.minimum(kq, Lq)
    # 64 k-means start points (as the headline search); node 0 beside it
    seeds = N.pick_seeds(ix, 64)
    res = {"workload": "uneven", "n": n, "nq": nq, "start_points": "%d k-means seeds" % len(seeds)}

    def batch(idx, fast, sp=seeds):
        """the batch's search captured by the library as one HIP graph
        (nbpg_search_uneven_graph_create): (replay, graph)"""
        qd = N.to_device(qpad[idx])
        Ld = torch.from_numpy(Lq[idx]).cuda()
        kd = torch.from_numpy(kq[idx]).cuda()
        Lmax, kmax = int(Lq[idx].max()), int(kq[idx].max())
        g = N.UnevenSearchGraph(ix, qd, Ld, kd, Lmax, kmax, fast=fast, start_points=sp)
        return g.replay, g

    # autotune the launch shape on the 4096-query batch
    idx4k = rng.choice(nq, 4096, replace=False)
    tune = []
    for wpg in (1, 2, 4):
        for W in (1, 2):
            for fl in (10, 11, 12):
                fast = (wpg, W, fl, 0, 0)
                try:
                    run, g = batch(idx4k, fast)
                    ms = events_ms(run, 5)
                    g.close()
                except N.NbpgError:
                    continue
                tune.append({"config": list(fast), "ms": round(ms, 4)})
    best = min(tune, key=lambda r: r["ms"])
    fast = tuple(best["config"])
    res["autotune"] = tune
    res["fast_config"] = list(fast)
    log("uneven autotune picked", fast)
    per_b = []
    for B, sp in ((1, seeds), (16, seeds), (256, seeds), (1024, seeds), (4096, seeds),
                  (4096, None)):
        idx = rng.choice(nq, B, replace=False)
        run, g = batch(idx, fast, sp)
        ms = events_ms(run, 20 if B <= 256 else 10)
        ids = run()[0].cpu().numpy().view(np.uint32)
        g.close()
        rec = []
        for j, qi in enumerate(idx):
            k = int(kq[qi])
            rec.append(recall(ids[j:j + 1, :k], gt[qi:qi + 1], gtd[qi:qi + 1], k))
        per_b.append({"batch": B, "start_points": "seeds" if sp is not None else "node 0",
                      "ms": round(ms, 4), "qps": round(B / ms * 1e3, 1),
                      "mean_recall@k": round(float(np.mean(rec)), 4)})
        log("uneven", per_b[-1])
    res["per_batch"] = per_b
    # batched RobustPrune: 100k nodes x 256 candidates, R 64, alpha (float)1.2
    P, C, Rr = 100_000, 256, 64
    nodes = rng.choice(n, P, replace=False).astype(np.uint32)
    cand = rng.integers(0, n, (P, C + 8)).astype(np.uint32)
    cands = np.zeros((P, C), np.uint32)
    for t in range(P):  # distinct, without the node itself
        u = np.unique(cand[t])
        u = u[u != nodes[t]][:C]
        if len(u) < C:
            extra = np.setdiff1d(rng.choice(n, 4 * C, replace=False), np.append(u, nodes[t]))
            u = np.concatenate([u, extra[:C - len(u)]])
        cands[t] = u
    # exact distances on the device
    cand_t = torch.from_numpy(cands.astype(np.int64)).cuda()
    node_t = torch.from_numpy(nodes.astype(np.int64)).cuda()
    bx = base_d.to(torch.int32)
    dist = torch.empty((P, C), dtype=torch.float32, device="cuda")
    for s in range(0, P, 4096):
        e = min(P, s + 4096)
        diff = bx[cand_t[s:e]] - bx[node_t[s:e]][:, None, :]
        dist[s:e] = (diff * diff).sum(-1).float()
    del bx
    dh = dist.cpu().numpy()
    rec = np.zeros((P, C), dtype=[("id", np.uint32), ("dist", np.float32)])
    rec["id"], rec["dist"] = cands, dh
    rec.sort(axis=1, order=["dist", "id"])
    packed = rec.reshape(-1).view(np.int64)
    cand_d = torch.from_numpy(packed.copy()).cuda()
    offs = torch.from_numpy(np.arange(P + 1, dtype=np.int64) * C).cuda()
    nodes_d = torch.from_numpy(nodes).cuda()
    alpha = float(np.float32(1.2))
    ms = events_ms(lambda: N.robust_prune_batch(ix, nodes_d, cand_d, offs, Rr, alpha=alpha), 3)
    gi, gd, gc = N.robust_prune_batch(ix, nodes_d, cand_d, offs, Rr, alpha=alpha)
    gi, gc = gi.cpu().numpy().view(np.uint32), gc.cpu().numpy()
    # bytes the prune reads per node: the candidates' rows and records
    pr = {"nodes": P, "candidates": C, "R": Rr, "ms": round(ms, 3),
          "nodes_per_s": round(P / ms * 1e3, 1),
          "GBs_candidate_rows": round(P * C * (base.shape[1] + 8) / (ms * 1e-3) / 1e9, 1)}
    log("batched prune %.3f ms" % ms)
    R, flags = ref_timing_lib()
    if R is not None and not args.no_cpu:
        c = rec.copy()
        cnt = np.zeros(P, np.uint32)
        cpu_s = R.nbpgr_prune_batch(0, base.ctypes.data, n, base.shape[1], nodes.ctypes.data, P,
                                    c.ctypes.data, C, Rr, alpha, 0.0, cnt.ctypes.data)
        same_cnt = cnt == gc.astype(np.uint32)
        col = np.arange(Rr)[None, :]
        live_mask = col < cnt[:, None]
        same_ids = ((gi == c["id"][:, :Rr]) | ~live_mask).all(axis=1)
        same = int((same_cnt & same_ids).sum())
        pr.update({"identical_lists": same, "all_identical": same == P,
                   "cpu_reference_s": round(cpu_s, 3), "cpu_threads": int(R.nbpgr_threads()),
                   "cpu_flags": flags, "speedup_vs_cpu": round(cpu_s * 1e3 / ms, 1)})
        log("prune identical %d / %d, reference %.2fs on %d threads" % (same, P, cpu_s,
                                                                         R.nbpgr_threads()))
    res["robust_prune"] = pr
    ix.close()
    return res


# ---------------------------------------------------------------------- serve
def run_serve(args, log):
    """10M x 96 i8, device Vamana build (R48, L96, alpha 1.2), then the
    continuous-batching server under Poisson request arrivals (1..64 queries
    per request, k = 10) at several offered loads: p50 / p99 latency, QPS and
    recall@10 of the served results."""
    torch = _torch()
    import nbpg_amd as N
    from nbpg_amd import datagen as D
    N.init(0)
    n, nq = args.n or 10_000_000, args.nq or 10_000
    t = time.time()
    base, q = D.workload_serve(n, nq)
    log("serve data %.1fs" % (time.time() - t))
    ix = N.Index(N.I8, n, base.shape[1], 48)
    ix.set_points(base)
    torch.cuda.synchronize()
    t = time.time()
    # the reference's Vamana flags --R 48 --L 96 --max_alpha 1.2, defaults for the
    # rest (num_passes 2, num_split 4, VL 140: Vamana.cpp:24-50)
    N.build(ix, 48, 96, VL=140, alpha=1.2, passes=2, num_split=4, sp_seed=1)
    torch.cuda.synchronize()
    build_s = time.time() - t
    log("serve device build %.1fs" % build_s)
    base_d = N.to_device(base.view(np.uint8))
    gt_ids, gt_d = N.compgt(base_d, torch.from_numpy(q.view(np.uint8)).cuda(), 100, N.I8,
                            N.GT_MFMA)
    gt, gtd = gt_ids.cpu().numpy().view(np.uint32), gt_d.cpu().numpy()
    del base_d
    q_d = N.to_device(pad_rows(q, ix.row_stride))
    fast = (2, 1, 11, 0, 0)
    # start points: 64 k-means seeds or the node-0 entry, whichever serves
    # more queries per second at recall@10 >= 0.95 on the whole batch (both
    # reported)
    seeds = N.pick_seeds(ix, 64)
    picks = {}
    for tag, sp in (("seeds", seeds), ("entry0", None)):
        for L in (16, 20, 24, 32, 40, 48, 64, 80, 96, 112, 128, 144, 160, 176, 192, 208, 224,
                  256):
            out = N.search(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=fast, start_points=sp)
            r = recall(out[0].cpu().numpy().view(np.uint32), gt, gtd, 10)
            log("serve %s L=%d recall %.4f" % (tag, L, r))
            if r >= 0.95:
                ms = replay_ms(N.SearchGraph(ix, q_d, L, 10, mode=N.SEARCH_FAST, fast=fast,
                                             start_points=sp), 10)
                picks[tag] = (L, r, nq / ms * 1e3)
                break
    if not picks:
        raise SystemExit("no beam width up to 256 reaches recall@10 0.95")
    use = max(picks, key=lambda t: picks[t][2])
    L_pick, rec_pick, full_qps = picks[use]
    sp_use = seeds if use == "seeds" else None
    log("serve %s L=%s recall %.4f, whole-batch %.0f QPS" % (use, L_pick, rec_pick, full_qps))
    res = {"workload": "serve", "n": n, "nq": nq, "device_build_s": round(build_s, 2),
           "start_points": ("%d k-means seeds" % len(seeds)) if use == "seeds" else "node 0",
           "L": L_pick, "batch_recall@10": round(rec_pick, 4),
           "whole_batch_qps": round(full_qps, 1),
           "start_point_choices": {t: {"L": v[0], "recall@10": round(v[1], 4),
                                       "whole_batch_qps": round(v[2], 1)}
                                   for t, v in picks.items()}}
    srv = N.Server(ix, slots=3, max_batch_queries=4096, mem_budget=256 << 20, max_L=256,
                   max_k=10, fast=fast, start_points=sp_use)
    loads = serve_loads(srv, q, gt, gtd, L_pick,
                        [0.01 * full_qps, 0.04 * full_qps, None, ("burst", 64)], 4.0, log)
    srv.close()
    res["loads"] = loads
    ix.close()
    return res


def serve_loads(srv, q, gt, gtd, L, targets, duration, log, k=10, seed=5):
    """Drive a running nbpg server with requests of 1..64 consecutive rows
    of q (k results each, beam width L) for `duration` seconds per target:
    Poisson arrivals at each offered QPS in `targets` (queries per second);
    None: a closed loop with 256 single requests outstanding; ("burst", B):
    a closed loop that submits bursts of B requests in one call
    (nbpg_server_submit_many) and keeps 8 bursts outstanding.  One row per
    target: offered / submitted / served QPS, the server's p50 / p99 / max
    request latency (submit to completion), mean batch size and recall@k of
    the first 100k served queries against (gt, gtd)."""
    rng = np.random.default_rng(seed)
    nq = q.shape[0]
    sizes = rng.integers(1, 65, 400_000)
    starts = rng.integers(0, nq - 64, 400_000)
    loads = []
    for target in targets:
        srv.stats(reset=True)
        burst = target[1] if isinstance(target, tuple) else 0
        gaps = rng.exponential(32.5 / (target if isinstance(target, float) else 1.0), 400_000)
        tickets = []
        done = threading.Event()
        results = {}
        window = threading.Semaphore(8 if burst else 256) if not isinstance(target, float) \
            else None

        def consumer():
            i, kept = 0, 0
            while True:
                while i >= len(tickets):
                    if done.is_set() and i >= len(tickets):
                        return
                    time.sleep(0.0002)
                tk, s0, m = tickets[i]
                if burst:
                    ids = srv.wait_many(tk)[0].reshape(m, k)
                else:
                    ids = srv.wait(tk)[0]
                if window is not None:
                    window.release()
                # ids kept for the recall of the first 100k served queries
                results[i] = (s0, m, ids if kept < 100_000 else None)
                kept += m
                i += 1

        th = threading.Thread(target=consumer, daemon=True)
        th.start()
        t0 = time.perf_counter()
        nxt = t0
        i = 0
        try:
            while i < len(gaps):
                if window is not None:
                    window.acquire()
                    if time.perf_counter() - t0 > duration:
                        break
                else:
                    nxt += gaps[i]
                    if nxt - t0 > duration:
                        break
                    now = time.perf_counter()
                    if nxt > now:
                        time.sleep(max(0.0, nxt - now - 0.0002))
                        while time.perf_counter() < nxt:
                            pass
                if burst:
                    sz = sizes[(i * burst) % (len(sizes) - burst):][:burst]
                    m = int(sz.sum())
                    s0 = int(starts[i]) % (nq - m + 1)
                    tickets.append((srv.submit_many(q[s0:s0 + m], sz, L, k), s0, m))
                else:
                    s0, m = int(starts[i]), int(sizes[i])
                    tickets.append((srv.submit(q[s0:s0 + m], L, k), s0, m))
                i += 1
        finally:
            submit_end = time.perf_counter()
            done.set()
        th.join(timeout=60)
        wall = time.perf_counter() - t0
        st = srv.stats()
        # recall over the first 100k served queries (all of them in the open loops)
        got = [j for j in range(len(tickets)) if j in results and results[j][2] is not None]
        ids = np.concatenate([results[j][2] for j in got])
        qi = np.concatenate([np.arange(results[j][0], results[j][0] + results[j][1]) for j in got])
        rec = recall(ids, gt[qi], gtd[qi], k)
        if isinstance(target, float):
            offered = round(target, 1)
        elif burst:
            offered = "closed loop, bursts of %d requests, 8 bursts outstanding" % burst
        else:
            offered = "closed loop, 256 requests outstanding"
        row = {"offered_qps": offered,
               "submitted_qps": round(st["queries"] / (submit_end - t0), 1),
               "served_qps": round(st["queries"] / wall, 1), "requests": st["requests"],
               "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
               "max_ms": round(st["max_ms"], 3),
               "mean_batch_queries": round(st["mean_batch_queries"], 1),
               "recall@%d" % k: round(rec, 4)}
        loads.append(row)
        log("serve", row)
    return loads
