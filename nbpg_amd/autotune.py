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
    "waves_per_group": (2, 4),
    "expand_width": (1, 2),
    "filter_log2": (10, 11, 12),
    "row_split": (0, 1),
    "waves_per_cu": (16, 24, 0),  # 0: as many as LDS and registers allow
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
         log=None, start_points=None):
    """Returns (best, table): best = dict(config, L, ms, qps, recall)."""
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
                ids, _, _ = N.search(index, q_dev, L, k, mode=N.SEARCH_FAST, fast=cfg, stats=False,
                                     start_points=start_points)
                torch.cuda.synchronize()
                r = recall_with_ties(ids.cpu().numpy().view(np.uint32), gt_ids, gt_dists, k)
                if r >= target:
                    chosen = (L, r)
                    break
            if chosen is None:
                continue
            L, r = chosen
            out = N.search(index, q_dev, L, k, mode=N.SEARCH_FAST, fast=cfg, stats=False,
                           start_points=start_points)
            ms = _time(lambda: N.search(index, q_dev, L, k, mode=N.SEARCH_FAST, fast=cfg,
                                        out=out, start_points=start_points), reps)
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
