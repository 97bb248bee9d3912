"""Autotuner for the fast (persistent) search kernel.

Sweeps the kernel's launch knobs -- queries (wavefronts) per workgroup,
beam entries expanded per step, LDS visited-filter size -- and for each
shape the smallest beam width L that reaches the recall target on a
calibration set with known ground truth; keeps the shape with the highest
measured throughput.  All timing is device-side (HIP events on the launch
stream) over whole batches.
"""
import itertools
import time

import numpy as np

import nbpg_amd as N

DEFAULT_SPACE = {
    "waves_per_group": (2, 4, 8),
    "expand_width": (1, 2, 4),
    "filter_log2": (9, 10, 11),
}
DEFAULT_L = (10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80, 96, 128, 160, 200, 256)


def recall_with_ties(ids, gt_ids, gt_dists, k=10):
    """recall@k counting ground-truth ties at the k-th distance as hits
    (the reference's calRecall, check_nn_recall.h:126-167)."""
    hit = 0
    for i in range(len(ids)):
        g = gt_ids[i]
        n = k
        last = gt_dists[i, k - 1]
        while n < len(g) and gt_dists[i, n] == last:
            n += 1
        rs = set(ids[i, :k].tolist())
        hit += sum(1 for x in g[:n].tolist() if x in rs)
    return hit / (k * len(ids))


def _time(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune(index, q_dev, gt_ids, gt_dists, k=10, target=0.95, space=None, Ls=DEFAULT_L, reps=5,
         log=None):
    """Returns (best, table): best = dict(config, L, ms, qps, recall)."""
    import torch
    space = space or DEFAULT_SPACE
    nq = q_dev.shape[0]
    table = []
    best = None
    for wpg, w, fl in itertools.product(space["waves_per_group"], space["expand_width"],
                                        space["filter_log2"]):
        cfg = (wpg, w, fl, 0)
        try:
            chosen = None
            for L in Ls:
                if L < k:
                    continue
                ids, _, _ = N.search(index, q_dev, L, k, mode=N.SEARCH_FAST, fast=cfg, stats=False)
                torch.cuda.synchronize()
                r = recall_with_ties(ids.cpu().numpy().view(np.uint32), gt_ids, gt_dists, k)
                if r >= target:
                    chosen = (L, r)
                    break
            if chosen is None:
                continue
            L, r = chosen
            out = N.search(index, q_dev, L, k, mode=N.SEARCH_FAST, fast=cfg, stats=False)
            ms = _time(lambda: N.search(index, q_dev, L, k, mode=N.SEARCH_FAST, fast=cfg,
                                        out=out), reps)
        except N.NbpgError as e:  # shape not supported (LDS budget, ...)
            if log:
                log("autotune skip %s: %s" % (cfg, e))
            continue
        row = {"config": cfg, "L": L, "recall": r, "ms": ms, "qps": nq / (ms * 1e-3)}
        table.append(row)
        if log:
            log("autotune %s L=%d recall %.4f %.3f ms %.0f QPS" % (cfg, L, r, ms, row["qps"]))
        if best is None or row["qps"] > best["qps"]:
            best = row
    return best, table
