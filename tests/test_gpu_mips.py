"""The inner-product metric (--use_mips, DistanceMetric::MIPS) on the device:
distances, ground truth on every engine, reference-mode search, the build,
RobustPrune and the raw-row rerank / filtering search, against outputs
recorded from the compiled reference and the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TYPES = [("u8", "U8", O.U8), ("i8", "I8", O.I8)]


def live(rows):
    out = rows.copy()
    out[:, 1:][np.arange(rows.shape[1] - 1)[None, :] >= rows[:, :1]] = 0
    return out


def mips_index(N, base, dtype, R, rows=None):
    ix = N.Index(dtype, base.shape[0], base.shape[1], R)
    ix.set_metric(N.METRIC_MIPS)
    ix.set_points(base)
    if rows is not None:
        ix.set_graph(rows)
    return ix


def qdev(N, ix, q):
    pad = np.zeros((len(q), ix.row_stride), np.uint8)
    pad[:, :q.shape[1]] = q.view(np.uint8)
    return N.to_device(pad)


@pytest.mark.parametrize("name,nt,ot", TYPES)
def test_mips_distances_and_ground_truth(gpu, golden_mips, name, nt, ot):
    N, z = gpu, golden_mips
    import torch
    base, q = z[name + "_base"], z[name + "_queries"]
    ix = mips_index(N, base, getattr(N, nt), 24)
    ids = np.arange(0, len(base), 7, dtype=np.uint32)
    d = N.l2_batch(ix, N.to_device(q[3]), N.to_device(ids)).cpu().numpy()
    want = np.array([O.distance(ot | O.MIPS, base[i], q[3]) for i in ids], np.float32)
    assert (d == want).all()
    bd, qd = torch.from_numpy(base).cuda(), torch.from_numpy(q).cuda()
    # reference slot rule: the reference binary's .gt on both engines
    for eng in (N.GT_PLAIN, N.GT_MFMA):
        gi, gd = N.compgt(bd, qd, 50, getattr(N, nt), eng | N.GT_REFTIES | N.GT_MIPS)
        assert (gi.cpu().numpy().view(np.uint32) == z[name + "_gt_ids"]).all(), eng
        assert (gd.cpu().numpy() == z[name + "_gt_dists"]).all(), eng
    # exact (dist, id) top-k on both engines == the oracle's
    oi, od = O.compgt(base, q, 50, ot | O.MIPS, tie_mode=1)
    for eng in (N.GT_PLAIN, N.GT_MFMA):
        gi, gd = N.compgt(bd, qd, 50, getattr(N, nt), eng | N.GT_MIPS)
        assert (gi.cpu().numpy().view(np.uint32) == oi).all(), eng
        assert (gd.cpu().numpy() == od).all(), eng
    ix.close()


@pytest.mark.parametrize("name,nt,ot", TYPES)
def test_mips_search_build_prune_match_reference(gpu, golden_mips, name, nt, ot):
    N, z = gpu, golden_mips
    base, q = z[name + "_base"], z[name + "_queries"]
    dt = getattr(N, nt)
    g1 = z[name + "_graph1"]
    ix = mips_index(N, base, dt, 24, g1)
    qd = qdev(N, ix, q)
    for i, L in enumerate(z["L"]):
        ids, d, st = N.search(ix, qd, int(L), 10, mode=N.SEARCH_REF)
        assert (ids.cpu().numpy().view(np.uint32) == z[name + "_search_ids"][i]).all(), int(L)
        assert (st.cpu().numpy().view(np.uint32) == z[name + "_search_stats"][i]).all(), int(L)
        # FAST (one lane per row, then 16-byte pieces per lane): valid ids,
        # each with its own inner-product distance
        for cfg in (None, (2, 1, 10, 0, 1)):
            fi, fd, _ = N.search(ix, qd, int(L), 10, mode=N.SEARCH_FAST, fast=cfg)
            fi, fd = fi.cpu().numpy().view(np.uint32), fd.cpu().numpy()
            for r in range(0, len(q), 9):
                for j in range(10):
                    if fi[r, j] != 0xFFFFFFFF:
                        assert fd[r, j] == O.distance(ot | O.MIPS, base[fi[r, j]], q[r]), cfg
                live_d = fd[r][fi[r] != 0xFFFFFFFF]
                assert (np.diff(live_d) >= 0).all()
                assert len(set(fi[r][fi[r] != 0xFFFFFFFF].tolist())) == len(live_d)
    ix.close()
    ix = mips_index(N, base, dt, 24)
    N.build(ix, 24, 48, alpha=1.2, passes=1, num_split=0)
    assert (live(ix.get_graph()) == live(g1)).all()
    ix.close()
    ix = mips_index(N, base, dt, 24)
    N.build(ix, 24, 40, VL=60, alpha=1.1, min_alpha=1.05, passes=2, num_split=1,
            dynamic_prune=True)
    assert (live(ix.get_graph()) == live(z[name + "_graph2"])).all()
    ix.close()
    ix = mips_index(N, base, dt, 24)
    cand = N.to_device(N.pack_cands(z[name + "_prune_ids"], z[name + "_prune_dists"]))
    offs = N.to_device(z[name + "_prune_offsets"].astype(np.int64))
    nodes = N.to_device(z[name + "_prune_nodes"])
    alphas = N.to_device(z[name + "_prune_alpha"].astype(np.float32))
    pids, _, cnt = N.robust_prune_batch(ix, nodes, cand, offs, 16, alphas_dev=alphas)
    pids, cnt = pids.cpu().numpy().view(np.uint32), cnt.cpu().numpy()
    for t in range(len(z[name + "_prune_nodes"])):
        m = int(z[name + "_prune_cnt"][t])
        assert cnt[t] == m and (pids[t, :m] == z[name + "_prune_out"][t, :m]).all(), t
    ix.close()


def test_mips_rerank_and_filtering_match_reference(gpu, golden_mips):
    N, z = gpu, golden_mips
    import torch
    ix = mips_index(N, z["ex_codes"], N.U8, 16, z["ex_graph"])
    qd = qdev(N, ix, z["ex_qcodes"])
    xb, xq = torch.from_numpy(z["ex_raw"]).cuda(), torch.from_numpy(z["ex_rawq"]).cuda()
    for i, (L, k, rf) in enumerate(z["ex_settings"]):
        L, k, rf = int(L), int(k), int(rf)
        ids, d, st = N.search_rerank(ix, qd, xb, xq, L, k, rf, mode=N.SEARCH_REF)
        assert (ids.cpu().numpy().view(np.uint32) == z["ex_ids"][i][:, :k]).all(), (L, k, rf)
        assert (st.cpu().numpy().view(np.uint32) == z["ex_stats"][i]).all(), (L, k, rf)
        _, od, _ = O.search_exact(z["ex_graph"], z["ex_codes"], z["ex_qcodes"], O.U8 | O.MIPS,
                                  z["ex_raw"], z["ex_rawq"], O.F32 | O.MIPS, L, k, rf)
        assert (d.cpu().numpy() == od).all(), (L, k, rf)
        if rf:
            # the stand-alone rerank of the whole frontier agrees
            fr, _, _ = N.search(ix, qd, L, k, mode=N.SEARCH_REF, frontier=True)
            a, ad = N.rerank(xb, xq, fr, k, rf, metric=N.METRIC_MIPS)
            assert (a == ids).all() and (ad == d).all()
    ix.close()
