"""nbpg_pick_seeds: k-means start nodes computed on the device, against a
numpy restatement of the same steps (FP64, dimension order, integer centre
sums: bit-identical), for every row type; and their use as FAST start
points."""
import numpy as np
import pytest

import oracle as O
from nbpg_amd import datagen as D

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _mix(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def _values(rows, dtype):
    """rows as the distance sees them: u8 / i8 values, packed bytes as the
    two 16-step nibble values (high nibble first)"""
    if dtype == O.I8:
        return rows.view(np.int8).astype(np.float64)
    if dtype in (O.U4X2, O.I4X2):
        r = rows.astype(np.int64)
        return np.stack([r & 0xF0, (r << 4) & 0xFF], axis=2).reshape(len(rows), -1).astype(
            np.float64)
    return rows.astype(np.float64)


def restate(base, dtype, k, S, iters, seed):
    n = len(base)
    ids = np.array([_mix(seed ^ (i << 20) ^ i) % n for i in range(S)], np.int64)
    x = _values(base[ids], dtype)
    c = x[:k].copy()
    for _ in range(iters):
        acc = np.zeros((S, k))
        for d in range(x.shape[1]):
            v = x[:, None, d] - c[None, :, d]
            acc = acc + v * v
        lab = acc.argmin(axis=1)
        for j in range(k):
            m = lab == j
            if m.any():
                c[j] = x[m].sum(axis=0) / m.sum()
    acc = np.zeros((S, k))
    for d in range(x.shape[1]):
        v = x[:, None, d] - c[None, :, d]
        acc = acc + v * v
    near = ids[acc.argmin(axis=0)]
    return np.unique(near).astype(np.uint32)


@pytest.mark.parametrize("dt,dims", [("u8", 64), ("i8", 48), ("u4", 32)])
def test_pick_seeds_matches_restatement(gpu, dt, dims):
    N = gpu
    t = {"u8": N.U8, "i8": N.I8, "u4": N.U4X2}[dt]
    if dt == "i8":
        base = D.mixture_i8(5000, dims, 20, 30.0, seed=3).view(np.uint8)
    else:
        base = D.mixture_u8(5000, dims, 20, 25.0, seed=3)
    ix = N.Index(t, len(base), dims, 16)
    ix.set_points(base)
    for k, S, iters, seed in [(16, 2000, 5, 0), (7, 500, 0, 11), (32, 3000, 3, 5)]:
        got = N.pick_seeds(ix, k, S, iters, seed)
        want = restate(base, {"u8": O.U8, "i8": O.I8, "u4": O.U4X2}[dt], k, S, iters, seed)
        assert (got == want).all(), (dt, k, S, iters, seed)
        assert (N.pick_seeds(ix, k, S, iters, seed) == got).all()
    with pytest.raises(N.NbpgError):
        N.pick_seeds(ix, 0)
    ix.close()


def test_seeds_improve_fast_search(gpu):
    """64 seeds from the picker: FAST recall@10 at a small beam is at least
    that of the single entry node 0, on 20k x 128 u8 with a built graph."""
    N = gpu
    base, q = D.workload_small(20000, 300)
    ix = N.Index(N.U8, len(base), 128, 32)
    ix.set_points(base)
    N.build(ix, 32, 64, alpha=1.2, passes=1, num_split=0)
    seeds = N.pick_seeds(ix, 64)
    assert 1 < len(seeds) <= 64 and (seeds < len(base)).all()
    assert (np.diff(seeds.astype(np.int64)) > 0).all()
    qp = np.zeros((len(q), ix.row_stride), np.uint8)
    qp[:, :128] = q
    qd = N.to_device(qp)
    gt, gtd = O.compgt(base, q, 100, O.U8, tie_mode=1)
    from nbpg_amd.autotune import recall_with_ties
    r0 = recall_with_ties(N.search(ix, qd, 16, 10, mode=N.SEARCH_FAST)[0].cpu().numpy(), gt,
                          gtd, 10)
    rs = recall_with_ties(N.search(ix, qd, 16, 10, mode=N.SEARCH_FAST,
                                   start_points=seeds)[0].cpu().numpy(), gt, gtd, 10)
    assert rs >= r0, (rs, r0)
    ix.close()
