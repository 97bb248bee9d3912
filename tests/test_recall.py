"""recall@k with ground-truth ties (the headline's recall figure) against the
oracle's restatement of calRecall (check_nn_recall.h:126-167) and a direct
count on tie-heavy data."""
import numpy as np

import oracle as O
from nbpg_amd.autotune import recall_with_ties


def _tie_heavy(seed, n=3000, nq=60, dims=8):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 3, (n, dims), dtype=np.uint8)
    q = rng.integers(0, 3, (nq, dims), dtype=np.uint8)
    return base, q


def _direct(ids, base, q, gt, k):
    """Pure-Python count: every ground-truth id of rank < k, and of any
    later rank whose distance continues the run equal to the k-th, that
    appears among the first k reported ids."""
    hits = 0
    for i in range(len(q)):
        d = lambda j: int(((base[j].astype(np.int64) - q[i]) ** 2).sum())  # noqa: E731
        last = d(gt[i, k - 1])
        adm = list(gt[i, :k])
        for j in gt[i, k:]:
            if d(j) != last:
                break
            adm.append(j)
        rep = set(int(x) for x in ids[i, :k])
        hits += sum(1 for j in adm if int(j) in rep)
    return hits / (k * len(q))


def test_recall_with_ties_matches_oracle():
    for seed in range(3):
        base, q = _tie_heavy(seed)
        gt, gtd = O.compgt(base, q, 100, O.U8, tie_mode=1)
        # each query has long runs of equal distances at rank 10
        assert (gtd[:, 10] == gtd[:, 9]).mean() > 0.5
        rng = np.random.default_rng(100 + seed)
        for k in (1, 10):
            ids = np.empty((len(q), k), np.uint32)
            for i in range(len(q)):
                pool = np.concatenate([gt[i, :3 * k], rng.integers(0, len(base), 2 * k)])
                ids[i] = rng.choice(pool, k, replace=False)
            ids[::7, -1] = 0xFFFFFFFF  # short result lists
            want = O.recall(ids, gt, base, q, O.U8, k)
            got = recall_with_ties(ids, gt, gtd, k)
            assert abs(got - want) < 1e-6, (seed, k, got, want)
            assert abs(got - _direct(ids, base, q, gt, k)) < 1e-12


def test_recall_with_ties_exact_results():
    base, q = _tie_heavy(7)
    gt, gtd = O.compgt(base, q, 50, O.U8, tie_mode=1)
    assert recall_with_ties(gt[:, :10], gt, gtd, 10) == 1.0
    # a tied id from past rank k counts as a hit in place of the one it ties
    ids = gt[:, :10].copy()
    swap = gtd[:, 10] == gtd[:, 9]
    ids[swap, 9] = gt[swap, 10]
    assert recall_with_ties(ids, gt, gtd, 10) == 1.0
    assert abs(O.recall(ids, gt, base, q, O.U8, 10) - 1.0) < 1e-6
