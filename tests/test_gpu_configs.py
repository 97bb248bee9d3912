"""Each benchmark configuration's own workload through the HIP path, checked
against the oracle (pinned by the reference-recorded fixtures), the
reference's own compiled prune where it is built, and closed forms:

* 100k x 128 u8, one-pass Vamana (R32 L75 a1.2): device build == oracle
  build row for row; reference-mode search ids and {visited, dist_cmps}
  equal the oracle's at L = 16..128 for 1000 queries.
* 1M x 128 u8 heavy-tailed mixture, NBPG graph (R64, 2 passes, 4 splits,
  dynamic prune): int8-MFMA top-100 == VALU top-100 for 10k queries and ==
  the oracle's exact top-100 for 256 of them; reference-mode search ==
  oracle; FAST search from 64 k-means seeds: exact distances, recall@10 >=
  0.95 at L = 16.
* Uneven batches (1, 16, 256, 4096 queries, L log-uniform 16..256, k in
  {1, 10, 100}) on that graph, and batched RobustPrune of 100k nodes x 256
  candidates (R64, alpha (float)1.2) == the reference's prune.
* 1M x 960 f32 decaying-spectrum pipeline at 200k rows: FP64 statistics
  against numpy, the bf16 projection within its derived bound of FP64,
  codes == the quantizer of the kernel's y, R48 builds and FAST search on
  the i8 and u4x2 codes with exact distances.
* 96-dim i8 serving at 1M rows (4096 clusters, sigma 30, R48 L96 a1.2
  device build): served results alone == in a mix, exact distances,
  recall@10 >= 0.95.
"""
import threading

import numpy as np
import pytest

import oracle as O
from nbpg_amd import datagen as D

pytestmark = pytest.mark.gpu


def pad(a, stride):
    out = np.zeros((a.shape[0], stride), np.uint8)
    out[:, :a.shape[1]] = a.view(np.uint8)
    return out


def exact_dists(base, q, ids, signed=False):
    """closed-form squared L2 of each returned id (int64), nq x k"""
    b = base.view(np.int8) if signed else base
    qq = q.view(np.int8) if signed else q
    d = b[ids].astype(np.int64) - qq[:, None, :].astype(np.int64)
    return (d * d).sum(-1)


def check_results(base, q, ids, d, dtype, k=None, signed=False, oracle_rows=50):
    """sorted, distinct, exact distances (closed form for all, the oracle's
    distance for the first rows)"""
    k = ids.shape[1] if k is None else k
    ids, d = ids[:, :k], d[:, :k]
    assert (ids < base.shape[0]).all()
    assert (np.diff(d, axis=1) >= 0).all()
    s = np.sort(ids, axis=1)
    assert (s[:, 1:] != s[:, :-1]).all()
    if dtype in (O.U8, O.I8):
        assert (exact_dists(base, q, ids, signed) == d.astype(np.int64)).all()
    for j in range(min(oracle_rows, len(q))):
        ref = np.array([O.distance(dtype, base[i], q[j]) for i in ids[j]], np.float32)
        assert (ref == d[j]).all(), j


def recall(ids, gt, gtd, k):
    from nbpg_amd.autotune import recall_with_ties
    return recall_with_ties(ids, gt, gtd, k)


# ------------------------------------------------------------ 100k x 128 u8


def test_c1_build_and_reference_search(gpu):
    N = gpu
    base, q = D.workload_small(100_000, 1000)
    ix = N.Index(N.U8, base.shape[0], base.shape[1], 32)
    ix.set_points(base)
    iso = N.build(ix, 32, 75, alpha=1.2, passes=1, num_split=0)
    rows, oiso = O.build_graph(base, O.U8, 32, 75, alpha=1.2, passes=1, num_split=0)
    assert iso == oiso
    g = ix.get_graph()
    live = np.arange(32)[None, :] < g[:, :1]
    assert (g[:, 0] == rows[:, 0]).all()
    assert ((g[:, 1:] == rows[:, 1:]) | ~live).all()
    qd = N.to_device(pad(q, ix.row_stride))
    for L in (16, 32, 64, 128):
        ids, d, st = N.search(ix, qd, L, 10, mode=N.SEARCH_REF)
        oi, od, ost = O.search(rows, base, q, O.U8, L, 10)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), L
        assert (d.cpu().numpy() == od).all(), L
        assert (st.cpu().numpy().view(np.uint32) == ost).all(), L
    ix.close()


# ------------------------------------------------------------ 1M x 128 u8


@pytest.fixture(scope="module")
def c2(gpu):
    N = gpu
    import torch
    base, q = D.workload_1m(1_000_000, 10_000)
    ix = N.Index(N.U8, base.shape[0], base.shape[1], 64)
    ix.set_points(base)
    N.build(ix, 64, 70, VL=140, alpha=1.1, min_alpha=1.05, passes=2, num_split=4,
            dynamic_prune=True, sp_seed=1)
    base_d = N.to_device(base)
    q_gt = torch.from_numpy(q).cuda()
    gi, gd = N.compgt(base_d, q_gt, 100, N.U8, N.GT_MFMA)
    env = dict(N=N, base=base, q=q, ix=ix, base_d=base_d, q_gt=q_gt,
               qd=N.to_device(pad(q, ix.row_stride)), gt=gi.cpu().numpy().view(np.uint32),
               gtd=gd.cpu().numpy())
    yield env
    ix.close()


def test_c2_groundtruth_mfma(c2):
    """int8-MFMA top-100 over the full 1M rows: == the VALU engine for all
    10k queries, == the oracle's exact (dist, id) top-100 for 256."""
    N = c2["N"]
    pi, pd = N.compgt(c2["base_d"], c2["q_gt"], 100, N.U8, N.GT_PLAIN)
    assert (pi.cpu().numpy().view(np.uint32) == c2["gt"]).all()
    assert (pd.cpu().numpy() == c2["gtd"]).all()
    oi, od = O.compgt(c2["base"], c2["q"][:256], 100, O.U8, tie_mode=1)
    assert (oi == c2["gt"][:256]).all()
    assert (od == c2["gtd"][:256]).all()


def test_c2_reference_search(c2):
    N, ix = c2["N"], c2["ix"]
    rows = ix.get_graph()
    q = c2["q"][:200]
    for L in (16, 64):
        ids, d, st = N.search(ix, c2["qd"][:200], L, 10, mode=N.SEARCH_REF)
        oi, od, ost = O.search(rows, c2["base"], q, O.U8, L, 10)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), L
        assert (d.cpu().numpy() == od).all(), L
        assert (st.cpu().numpy().view(np.uint32) == ost).all(), L


@pytest.mark.parametrize("fast,L", [((4, 1, 10, 0, 0), 16), ((2, 1, 11, 0, 0), 20)])
def test_c2_fast_search(c2, fast, L):
    """The headline search: 10k queries, 64 k-means seeds, one launch."""
    N = c2["N"]
    seeds = N.pick_seeds(c2["ix"], 64)
    ids, d, st = N.search(c2["ix"], c2["qd"], L, 10, mode=N.SEARCH_FAST, fast=fast,
                          start_points=seeds)
    ids, d = ids.cpu().numpy().view(np.uint32), d.cpu().numpy()
    check_results(c2["base"], c2["q"], ids, d, O.U8)
    r = recall(ids, c2["gt"], c2["gtd"], 10)
    assert r >= 0.95, r
    st = st.cpu().numpy()
    assert (st[:, 1] >= len(seeds)).all() and (st[:, 0] >= 1).all()


def test_c4_uneven_batches(c2):
    N = c2["N"]
    import torch
    rng = np.random.default_rng(17)
    nq = len(c2["q"])
    Lq = np.exp(rng.uniform(np.log(16), np.log(256), nq)).round().astype(np.int32)
    kq = np.minimum(rng.choice(np.array([1, 10, 100], np.int32), nq), Lq)
    start = 0
    for B in (1, 16, 256, 4096):
        idx = np.arange(start, start + B)
        start += B
        Ld = torch.from_numpy(Lq[idx]).cuda()
        kd = torch.from_numpy(kq[idx]).cuda()
        ids, d, _ = N.search_uneven(c2["ix"], c2["qd"][idx[0]:idx[-1] + 1], Ld, kd,
                                    int(Lq[idx].max()), int(kq[idx].max()),
                                    fast=(2, 1, 11, 0, 0))
        ids, d = ids.cpu().numpy().view(np.uint32), d.cpu().numpy()
        hits, want = 0.0, 0
        for j, qi in enumerate(idx):
            k = int(kq[qi])
            check_results(c2["base"], c2["q"][qi:qi + 1], ids[j:j + 1], d[j:j + 1], O.U8, k=k,
                          oracle_rows=1 if j < 20 else 0)
            hits += recall(ids[j:j + 1, :k], c2["gt"][qi:qi + 1], c2["gtd"][qi:qi + 1], k) * k
            want += k
        assert hits / want >= 0.95, (B, hits / want)


def test_c4_batched_prune_100k(c2):
    """100k nodes x 256 distinct candidates sorted by (dist, id), R 64,
    alpha (float)1.2: every list == the reference's prune (its compiled
    code, where built) and == the oracle's on a sample."""
    N = c2["N"]
    import torch
    base, n = c2["base"], c2["base"].shape[0]
    rng = np.random.default_rng(23)
    P, Cn, Rr = 100_000, 256, 64
    nodes = rng.choice(n, P, replace=False).astype(np.uint32)
    cand = np.empty((P, Cn), np.uint32)
    for t in range(P):
        u = np.unique(rng.integers(0, n, Cn + 16).astype(np.uint32))
        u = u[u != nodes[t]]
        while len(u) < Cn:
            u = np.union1d(u, rng.integers(0, n, Cn).astype(np.uint32))
            u = u[u != nodes[t]]
        cand[t] = rng.permutation(u)[:Cn]
    bx = c2["base_d"].to(torch.int32)
    ct = torch.from_numpy(cand.astype(np.int64)).cuda()
    nt = torch.from_numpy(nodes.astype(np.int64)).cuda()
    dist = torch.empty((P, Cn), dtype=torch.float32, device="cuda")
    for s in range(0, P, 4096):
        e = min(P, s + 4096)
        diff = bx[ct[s:e]] - bx[nt[s:e]][:, None, :]
        dist[s:e] = (diff * diff).sum(-1).float()
    del bx
    rec = np.zeros((P, Cn), dtype=[("id", np.uint32), ("dist", np.float32)])
    rec["id"], rec["dist"] = cand, dist.cpu().numpy()
    rec.sort(axis=1, order=["dist", "id"])
    cand_d = torch.from_numpy(rec.reshape(-1).view(np.int64).copy()).cuda()
    offs = torch.from_numpy(np.arange(P + 1, dtype=np.int64) * Cn).cuda()
    alpha = float(np.float32(1.2))
    gi, gd, gc = N.robust_prune_batch(c2["ix"], torch.from_numpy(nodes).cuda(), cand_d, offs, Rr,
                                      alpha=alpha)
    gi, gd, gc = gi.cpu().numpy().view(np.uint32), gd.cpu().numpy(), gc.cpu().numpy()
    live = np.arange(Rr)[None, :] < gc[:, None]
    R = O.ref()
    if R is not None:
        c = rec.copy()
        cnt = np.zeros(P, np.uint32)
        R.nbpgr_prune_batch(0, base.ctypes.data, n, base.shape[1], nodes.ctypes.data, P,
                            c.ctypes.data, Cn, Rr, alpha, 0.0, cnt.ctypes.data)
        assert (cnt == gc).all()
        assert ((gi == c["id"][:, :Rr]) | ~live).all()
        assert ((gd == c["dist"][:, :Rr]) | ~live).all()
    for t in rng.choice(P, 300, replace=False):
        oi, od = O.prune(base, O.U8, int(nodes[t]), rec["id"][t], rec["dist"][t], Rr, alpha)
        assert gc[t] == len(oi) and (gi[t, :len(oi)] == oi).all(), t
        assert (gd[t, :len(oi)] == od).all(), t


# ------------------------------------------------------------ 960-dim PCA


def test_c3_pca_pipeline(gpu):
    N = gpu
    import torch
    from nbpg_amd import pca as P
    x, xq = D.workload_pca(200_000, 2000)
    # FP64 statistics against numpy's float64 on the same rows
    mean, cov = N.pca_stats(x)
    xh = x.cpu().numpy().astype(np.float64)
    m64 = xh.mean(axis=0)
    xc = xh - m64
    c64 = (xc.T @ xc) / (len(xh) - 1)
    del xc
    assert np.abs(mean.cpu().numpy() - m64).max() <= 1e-12 * np.abs(m64).max() + 1e-15
    assert np.abs(cov.cpu().numpy() - c64).max() <= 1e-10 * np.abs(c64).max()
    f = P.fit(x, 64)
    W, b = f["W"], f["bias"]
    Wd, bd = torch.from_numpy(W).cuda(), torch.from_numpy(b).cuda()
    # bf16 projection within its derived bound of the FP64 projection (each
    # operand rounded to bf16, 2^-9 relative each; fp32 sums), on a sample
    sel = np.random.default_rng(3).choice(len(xh), 4000, replace=False)
    y64 = xh[sel] @ W.astype(np.float64).T + b
    bound = np.abs(xh[sel]) @ np.abs(W.astype(np.float64)).T * 2.0 ** -7 + 1e-6
    codes = {}
    for code, sg, ch in ((N.CODE_I8, 1, 1), (N.CODE_U4, 0, 2)):
        c, y = N.pca_project(x, Wd, bd, code=code, mode=N.PCA_BF16, want_y=True)
        yh = y.cpu().numpy()
        err = np.abs(yh[sel] - y64)
        assert (err <= bound).all(), err.max()
        assert err.max() < 2e-3 and err.mean() < 2e-4, (err.max(), err.mean())
        assert (c.cpu().numpy() == O.compress(yh, sg, ch)).all(), code
        cq, _ = N.pca_project(xq, Wd, bd, code=code, mode=N.PCA_BF16)
        codes[code] = (c.cpu().numpy(), cq.cpu().numpy())
    del xh
    # R48 builds and FAST search on the codes: exact integer distances
    for code, dt in ((N.CODE_I8, N.I8), (N.CODE_U4, N.U4X2)):
        cb, cq = codes[code]
        ix = N.Index(dt, cb.shape[0], cb.shape[1], 48)
        ix.set_points(cb)
        N.build(ix, 48, 96, alpha=1.2, passes=1, num_split=0)
        qd = N.to_device(pad(cq, ix.row_stride))
        ids, d, _ = N.search(ix, qd, 32, 10, mode=N.SEARCH_FAST)
        ids, d = ids.cpu().numpy().view(np.uint32), d.cpu().numpy()
        odt = O.I8 if dt == N.I8 else O.U4X2
        check_results(cb, cq, ids, d, odt, signed=dt == N.I8, oracle_rows=100)
        gt, gtd = O.compgt(cb, cq[:200], 100, odt, tie_mode=1)
        r = recall(ids[:200], gt, gtd, 10)
        assert r >= (0.9 if dt == N.I8 else 0.5), (dt, r)
        ix.close()


# ------------------------------------------------------------ 96-dim i8 serving


def test_c5_serving_i8(gpu):
    N = gpu
    import torch
    n, nq = 1_000_000, 2000
    base, q = D.workload_serve(n, nq)
    ix = N.Index(N.I8, n, base.shape[1], 48)
    ix.set_points(base)
    N.build(ix, 48, 96, VL=140, alpha=1.2, passes=2, num_split=4, sp_seed=1)
    bu = base.view(np.uint8)
    gi, gd = N.compgt(N.to_device(bu), torch.from_numpy(q.view(np.uint8)).cuda(), 100, N.I8,
                      N.GT_MFMA)
    gt, gtd = gi.cpu().numpy().view(np.uint32), gd.cpu().numpy()
    oi, od = O.compgt(bu, q.view(np.uint8)[:64], 100, O.I8, tie_mode=1)
    assert (oi == gt[:64]).all() and (od == gtd[:64]).all()
    fast = (2, 1, 11, 0, 0)
    qd = N.to_device(pad(q, ix.row_stride))
    L = None
    for Lc in (32, 48, 64, 96, 128, 160):
        ids, _, _ = N.search(ix, qd, Lc, 10, mode=N.SEARCH_FAST, fast=fast)
        if recall(ids.cpu().numpy().view(np.uint32), gt, gtd, 10) >= 0.95:
            L = Lc
            break
    assert L is not None
    rng = np.random.default_rng(5)
    reqs, off = [], 0
    while off < nq:
        m = int(min(rng.integers(1, 65), nq - off))
        reqs.append((off, m))
        off += m
    srv = N.Server(ix, slots=1, max_batch_queries=64, max_L=256, max_k=10, fast=fast)
    alone = []
    for s0, m in reqs[:60]:
        ids, d, _ = srv.wait(srv.submit(q[s0:s0 + m], L, 10))
        alone.append((ids, d))
    srv.close()
    srv = N.Server(ix, slots=3, max_batch_queries=1024, mem_budget=64 << 20, max_L=256,
                   max_k=10, fast=fast)
    tickets = [None] * len(reqs)

    def producer(part):
        for i in part:
            s0, m = reqs[i]
            tickets[i] = srv.submit(q[s0:s0 + m], L, 10)

    order = rng.permutation(len(reqs))
    th = [threading.Thread(target=producer, args=(order[j::4],)) for j in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    got = np.zeros((nq, 10), np.uint32)
    gdist = np.zeros((nq, 10), np.float32)
    for i, (s0, m) in enumerate(reqs):
        ids, d, _ = srv.wait(tickets[i])
        got[s0:s0 + m], gdist[s0:s0 + m] = ids.view(np.uint32), d
        if i < len(alone):
            assert (ids == alone[i][0]).all() and (d == alone[i][1]).all(), i
    st = srv.stats()
    srv.close()
    assert st["requests"] == len(reqs) and st["batches"] > 1
    check_results(base.view(np.uint8), q.view(np.uint8), got, gdist, O.I8, signed=True)
    assert recall(got, gt, gtd, 10) >= 0.95
    ix.close()
