"""Batched RobustPrune on every row type and size against the oracle's
Prune::sort_candidates + limitedSimilarityPrune (prune.h:49-74, 158-209).

Rows of up to 128 bytes (8-bit) or 64 bytes (packed 4-bit) and nodes with up
to 256 candidates take the matrix-core scan; longer rows and longer candidate
lists take the vector scan, both in the same call.  Data are clustered with
exact duplicates, so the alpha test meets equal and zero distances."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# (name, dims in bytes): 8-bit strides 32..128 (20 bytes pads to 32), 4-bit
# packed 32 / 64 bytes, and two sizes beyond the matrix-core path
TYPES = [("u8", 128), ("u8", 96), ("u8", 64), ("u8", 20), ("i8", 96), ("i8", 64),
         ("u4", 32), ("i4", 32), ("u4", 64), ("u8", 200), ("u4", 96)]


def clustered(rng, dt, n, dims):
    nc = 80
    if dt in ("u4", "i4"):
        cen = rng.integers(0, 16, (nc, 2 * dims))
        nib = np.clip(cen[rng.integers(0, nc, n)] + rng.integers(-2, 3, (n, 2 * dims)), 0, 15)
        rows = (nib[:, 0::2] | (nib[:, 1::2] << 4)).astype(np.uint8)
    else:
        lo, hi = (0, 255) if dt == "u8" else (-128, 127)
        cen = rng.integers(lo, hi + 1, (nc, dims))
        x = cen[rng.integers(0, nc, n)] + rng.normal(0, 18, (n, dims)).round().astype(np.int64)
        rows = np.clip(x, lo, hi).astype(np.uint8 if dt == "u8" else np.int8)
        rows[:4] = lo  # extreme rows
        rows[4:8] = hi
    rows[n // 2:n // 2 + 20] = rows[:20]  # exact duplicates
    return rows


def run_case(N, dt, dims, mips, alpha, cosa, dynamic, seed):
    rng = np.random.default_rng(seed)
    ntype = {"u8": N.U8, "i8": N.I8, "u4": N.U4X2, "i4": N.I4X2}[dt]
    otype = {"u8": O.U8, "i8": O.I8, "u4": O.U4X2, "i4": O.I4X2}[dt] | (O.MIPS if mips else 0)
    n = 1500
    base = clustered(rng, dt, n, dims)
    ix = N.Index(ntype, n, dims, 64)
    if mips:
        ix.set_metric(N.METRIC_MIPS)
    ix.set_points(base)
    P, R = 160, 48
    nodes = rng.integers(0, n, P).astype(np.uint32)
    offs = np.zeros(P + 1, np.int64)
    allids, alld = [], []
    for t in range(P):
        c = int(rng.choice([int(rng.integers(1, 40)), int(rng.integers(40, 257)),
                            int(rng.integers(257, 320))], p=[0.2, 0.65, 0.15]))
        ids = rng.integers(0, n, c).astype(np.uint32)  # duplicates included
        if t % 5 == 0:
            ids[c // 2] = nodes[t]
        allids.append(ids)
        alld.append(np.array([O.distance(otype, base[i], base[nodes[t]]) for i in ids],
                             np.float32))
        offs[t + 1] = offs[t] + c
    alphas = None
    if dynamic:
        alphas = rng.uniform(1.0, 1.5, P).astype(np.float32)
        alphas[::7] = 1.0
    cand = N.to_device(N.pack_cands(np.concatenate(allids), np.concatenate(alld)))
    gi, gd, gc = N.robust_prune_batch(ix, N.to_device(nodes), cand, N.to_device(offs), R,
                                      alpha=alpha, cos_angle=cosa,
                                      alphas_dev=None if alphas is None else N.to_device(alphas))
    gi, gd, gc = gi.cpu().numpy().view(np.uint32), gd.cpu().numpy(), gc.cpu().numpy()
    pruned_any = 0
    for t in range(P):
        a = float(alphas[t]) if alphas is not None else alpha
        oi, od = O.prune(base, otype, int(nodes[t]), allids[t], alld[t], R, a, cosa)
        assert gc[t] == len(oi), (t, gc[t], len(oi))
        assert (gi[t, :len(oi)] == oi).all(), t
        assert (gd[t, :len(oi)] == od).all(), t
        pruned_any += len(oi) < min(R, len(np.unique(allids[t])))
    ix.close()
    return pruned_any


@pytest.mark.parametrize("dt,dims", TYPES)
@pytest.mark.parametrize("mips", [0, 1])
def test_prune_types_match_oracle(gpu, dt, dims, mips):
    # inner-product distances are negative: alpha < 1 keeps more than the first
    pruned = run_case(gpu, dt, dims, mips, 0.9 if mips else 1.2, 0.0, False, 100 + dims + 7 * mips)
    assert pruned > 0  # the alpha test fired somewhere


@pytest.mark.parametrize("dt,dims", [("u8", 128), ("i8", 96), ("u4", 32)])
@pytest.mark.parametrize("alpha,cosa,dynamic", [(1.2, 0.5, False), (0.0, 0.3, False),
                                                (1.0, 0.0, True), (1.05, 0.0, False)])
def test_prune_settings_match_oracle(gpu, dt, dims, alpha, cosa, dynamic):
    run_case(gpu, dt, dims, 0, alpha, cosa, dynamic, 7 + dims)
