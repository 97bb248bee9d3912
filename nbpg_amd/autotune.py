"""Autotuner for the fast (persistent) search kernel.

Sweeps the kernel's launch knobs -- queries (wavefronts) per workgroup,
beam entries expanded per step, LDS visited-filter size, waves per CU --
and for each shape the smallest beam width L that reaches the recall
target on a calibration set with known ground truth; then times the shape
with consecutive batches issued round-robin over 1..4 streams (a batch's
last queries run alone on a mostly idle device; the next batch on another
stream fills it) and keeps the (shape, streams) pair with the highest
throughput.  Timing is over whole batches, between two device
synchronisations.
"""
import itertools
import math
import time

import numpy as np

import nbpg_amd as N

DEFAULT_SPACE = {
    "waves_per_group": (2, 4),
    "expand_width": (1, 2),
    "filter_log2": (10, 11, 12),
    "row_split": (0, 1),
    "waves_per_cu": (16, 24, 0),  # 0: as many as LDS and registers allow
    "streams": (1, 2, 3, 4),
}
DEFAULT_L = (10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80, 96, 128, 160, 200, 256)


def recall_with_ties(ids, gt_ids, gt_dists, k=10):
    """recall@k counting ground-truth ties at the k-th distance as hits
    (the reference's calRecall, check_nn_recall.h:126-167): ground-truth
    entries past k whose distance equals the k-th one are also admissible."""
    ids = np.asarray(ids)[:, :k].astype(np.int64)
    g = np.asarray(gt_ids).astype(np.int64)
    gd = np.asarray(gt_dists)
    last = gd[:, k - 1:k]
    admissible = np.zeros(g.shape, bool)
    admissible[:, :k] = True
    if g.shape[1] > k:
        # the ties must continue the run right after position k-1
        run = np.cumprod(gd[:, k:] == last, axis=1).astype(bool)
        admissible[:, k:] = run
    # a ground-truth id counts when it appears among the k reported ids
    hit = (g[:, :, None] == ids[:, None, :]).any(axis=2) & admissible
    return float(hit.sum()) / (k * len(ids))


def pipelined_ms(launch, streams, steps):
    """ms per step when `steps` steps are issued round-robin over `streams`
    streams: launch(i) issues step i on the current stream (stream i %
    streams).  Host clock between two device synchronisations."""
    import torch
    ss = [torch.cuda.Stream() for _ in range(streams)]

    def issue(n):
        for i in range(n):
            with torch.cuda.stream(ss[i % streams]):
                launch(i)
    torch.cuda.synchronize()
    issue(streams)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    issue(steps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def output_sets(batches, streams):
    """Output sets needed when step i searches batch i % batches on stream
    i % streams: steps i and i + lcm(batches, streams) share a set and run
    on one stream, so no two steps that may overlap write the same set."""
    return batches * streams // math.gcd(batches, streams)


def tune(index, q_dev, gt_ids, gt_dists, k=10, target=0.95, space=None, Ls=DEFAULT_L, steps=24,
         log=None, start_points=None, q_batches=None, mode=None):
    """Returns (best, table): best = dict(config, L, streams, ms, qps,
    recall); ms is per batch.  L is chosen on q_dev (with its ground truth);
    timing searches the batches of q_batches in turn (default: q_dev
    alone), each batch the size of q_dev.  mode: the FAST mode bits
    (default N.SEARCH_FAST; e.g. | N.SEARCH_SEED_NEAREST)."""
    mode = N.SEARCH_FAST if mode is None else mode
    import torch
    space = space or DEFAULT_SPACE
    nq = q_dev.shape[0]
    table = []
    best = None
    cu = N.device_info()["cu"]
    for wpg, w, fl, split, wpc in itertools.product(space["waves_per_group"], space["expand_width"],
                                                   space["filter_log2"], space["row_split"],
                                                   space["waves_per_cu"]):
        cfg = (wpg, w, fl, (cu * wpc) // wpg if wpc else 0, split)
        try:
            chosen = None
            for L in Ls:
                if L < k:
                    continue
                ids, _, _ = N.search(index, q_dev, L, k, mode=mode, fast=cfg, stats=False,
                                     start_points=start_points)
                torch.cuda.synchronize()
                r = recall_with_ties(ids.cpu().numpy().view(np.uint32), gt_ids, gt_dists, k)
                if r >= target:
                    chosen = (L, r)
                    break
            if chosen is None:
                continue
            L, r = chosen
            qb = q_batches or [q_dev]
            rows = []
            for st in space["streams"]:
                M = output_sets(len(qb), st)
                outs = [N.search(index, qb[j % len(qb)], L, k, mode=mode, fast=cfg,
                                 stats=False, start_points=start_points) for j in range(M)]

                def launch(i):
                    N.search(index, qb[i % len(qb)], L, k, mode=mode, fast=cfg,
                             out=outs[i % M], start_points=start_points)
                ms = pipelined_ms(launch, st, steps)
                rows.append({"config": cfg, "L": L, "streams": st, "recall": r, "ms": ms,
                             "qps": nq / (ms * 1e-3)})
        except N.NbpgError as e:  # shape not supported (LDS budget, ...)
            if log:
                log("autotune skip %s: %s" % (cfg, e))
            continue
        for row in rows:
            table.append(row)
            if log:
                log("autotune %s L=%d streams %d recall %.4f %.3f ms %.0f QPS"
                    % (cfg, L, row["streams"], r, row["ms"], row["qps"]))
            if best is None or row["qps"] > best["qps"]:
                best = row
    return best, table
