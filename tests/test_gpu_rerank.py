"""Rerank (beam_search_rerank, beamSearch.h:183-236): a FAST search over
quantized codes, then the frontier re-scored with full-precision rows, equals
the oracle's distances and (dist, id) order exactly."""
import numpy as np
import pytest

import oracle as O
from nbpg_amd import datagen as D

pytestmark = pytest.mark.gpu


def expected(raw, rq, cand, k, factor, dtype):
    out_i, out_d = [], []
    for j in range(len(rq)):
        c = cand[j]
        valid = np.nonzero(c == 0xFFFFFFFF)[0]
        m = valid[0] if len(valid) else len(c)
        m = min(m, k * factor)
        pts = sorted((O.distance(dtype, raw[i], rq[j]), int(i)) for i in c[:m])[:k]
        ids = [p[1] for p in pts] + [0xFFFFFFFF] * (k - len(pts))
        ds = [p[0] for p in pts] + [np.inf] * (k - len(pts))
        out_i.append(ids)
        out_d.append(ds)
    return np.array(out_i, np.uint32), np.array(out_d, np.float32)


@pytest.mark.parametrize("raw_kind", ["f32", "u8"])
def test_rerank_matches_oracle(gpu, raw_kind):
    N = gpu
    import torch
    rng = np.random.default_rng(12)
    if raw_kind == "f32":
        raw = (rng.standard_normal((6000, 96)) * np.linspace(2, 0.2, 96)).astype(np.float32)
        rq = (rng.standard_normal((150, 96)) * np.linspace(2, 0.2, 96)).astype(np.float32)
        codes = np.clip(np.rint(raw * 20 + 128), 0, 255).astype(np.uint8)
        qcodes = np.clip(np.rint(rq * 20 + 128), 0, 255).astype(np.uint8)
        ctype, rdt = N.U8, O.F32
    else:
        raw = D.mixture_u8(6000, 64, 30, 25.0, seed=3)
        rq = D.mixture_u8(150, 64, 30, 25.0, seed=4)
        codes = ((raw >> 4)[:, 0::2] << 4 | (raw >> 4)[:, 1::2]).astype(np.uint8)  # u4x2
        qcodes = ((rq >> 4)[:, 0::2] << 4 | (rq >> 4)[:, 1::2]).astype(np.uint8)
        ctype, rdt = N.U4X2, O.U8
    ix = N.Index(ctype, len(codes), codes.shape[1], 24)
    ix.set_points(codes)
    N.build(ix, 24, 48, alpha=1.2, passes=1, num_split=0)
    qpad = np.zeros((len(qcodes), ix.row_stride), np.uint8)
    qpad[:, :qcodes.shape[1]] = qcodes
    L = 40
    cand, _, _ = N.search(ix, N.to_device(qpad), L, L, mode=N.SEARCH_FAST)
    for k, factor in [(10, 3), (10, 100), (1, 5), (32, 1), (40, 1)]:
        ids, d = N.rerank(torch.from_numpy(raw).cuda(), torch.from_numpy(rq).cuda(), cand, k,
                          factor)
        ei, ed = expected(raw, rq, cand.cpu().numpy().view(np.uint32), k, factor, rdt)
        assert (ids.cpu().numpy().view(np.uint32) == ei).all(), (k, factor)
        assert (d.cpu().numpy() == ed).all(), (k, factor)
    # short frontiers (at most 32 candidates: several queries share a
    # wavefront), some ended early, k above the candidate count
    c16 = np.ascontiguousarray(cand.cpu().numpy().view(np.uint32)[:, :16])
    c16[5, 9:] = 0xFFFFFFFF
    c16[6, :] = 0xFFFFFFFF
    c16[149, 3:] = 0xFFFFFFFF
    cd16 = torch.from_numpy(c16.view(np.int32)).cuda()
    for k, factor in [(10, 1), (10, 2), (4, 1), (16, 1), (20, 1), (1, 1), (3, 3)]:
        ids, d = N.rerank(torch.from_numpy(raw).cuda(), torch.from_numpy(rq).cuda(), cd16, k,
                          factor)
        ei, ed = expected(raw, rq, c16, k, factor, rdt)
        assert (ids.cpu().numpy().view(np.uint32) == ei).all(), (k, factor)
        assert (d.cpu().numpy() == ed).all(), (k, factor)
    # the rerank can only improve the raw-space recall of the frontier head
    gt, _ = O.compgt(raw, rq, 10, rdt, tie_mode=1) if raw_kind == "u8" else (None, None)
    if gt is not None:
        ids, _ = N.rerank(torch.from_numpy(raw).cuda(), torch.from_numpy(rq).cuda(), cand, 10, 4)
        r_re = O.recall(ids.cpu().numpy().view(np.uint32), gt, raw, rq, O.U8, 10)
        r_q = O.recall(cand.cpu().numpy().view(np.uint32)[:, :10], gt, raw, rq, O.U8, 10)
        assert r_re >= r_q
    with pytest.raises(N.NbpgError):
        N.rerank(torch.from_numpy(raw).cuda(), torch.from_numpy(rq).cuda(),
                 torch.zeros((len(rq), 1100), dtype=torch.int32, device="cuda"), 100, 11)
    # long candidate lists (up to the 1024 limit), some ended early
    rng = np.random.default_rng(5)
    cand = np.stack([rng.permutation(len(raw))[:1024] for _ in range(len(rq))]).astype(np.uint32)
    cand[3, 500:] = 0xFFFFFFFF
    cand[7, 0:] = 0xFFFFFFFF
    cd = torch.from_numpy(cand.view(np.int32)).cuda()
    for k, factor in [(50, 14), (100, 11), (7, 1000)]:
        ids, d = N.rerank(torch.from_numpy(raw).cuda(), torch.from_numpy(rq).cuda(), cd, k, factor)
        ei, ed = expected(raw, rq, cand, k, factor, rdt)
        assert (ids.cpu().numpy().view(np.uint32) == ei).all(), (k, factor)
        assert (d.cpu().numpy() == ed).all(), (k, factor)
    # rows whose size is not a multiple of 16 bytes take the lane-per-row kernel
    odd, oq = np.ascontiguousarray(raw[:, :61]), np.ascontiguousarray(rq[:, :61])
    ids, d = N.rerank(torch.from_numpy(odd).cuda(), torch.from_numpy(oq).cuda(), cd[:, :300], 30, 10)
    ei, ed = expected(odd, oq, cand[:, :300], 30, 10, rdt)
    assert (ids.cpu().numpy().view(np.uint32) == ei).all()
    assert (d.cpu().numpy() == ed).all()
    ix.close()


WORKLOADS = [("f32", "U8", "F32"), ("u8", "U4X2", "U8")]


def _exact_index(N, z, name, cdt):
    import torch
    codes = z[name + "_codes"]
    ix = N.Index(getattr(N, cdt), codes.shape[0], codes.shape[1], 16)
    ix.set_points(codes)
    ix.set_graph(z[name + "_graph"])
    qc = z[name + "_qcodes"]
    qpad = np.zeros((len(qc), ix.row_stride), np.uint8)
    qpad[:, :qc.shape[1]] = qc
    return (ix, N.to_device(qpad), torch.from_numpy(z[name + "_raw"]).cuda(),
            torch.from_numpy(z[name + "_rawq"]).cuda())


@pytest.mark.parametrize("name,cdt,xdt", WORKLOADS)
def test_search_rerank_ref_matches_reference(gpu, golden_exact, name, cdt, xdt):
    """Reference mode with raw rows: rerank (rerank_factor > 0) and filtering
    (0) return the reference's ids and {visited, dist_cmps}, with the raw
    rows' distances."""
    N, z = gpu, golden_exact
    ix, qd, xb, xq = _exact_index(N, z, name, cdt)
    for i, (L, k, rf) in enumerate(z["settings"]):
        L, k, rf = int(L), int(k), int(rf)
        ids, d, st = N.search_rerank(ix, qd, xb, xq, L, k, rf, mode=N.SEARCH_REF)
        assert (ids.cpu().numpy().view(np.uint32) == z[name + "_ids"][i][:, :k]).all(), (L, k, rf)
        assert (st.cpu().numpy().view(np.uint32) == z[name + "_stats"][i]).all(), (L, k, rf)
        _, od, _ = O.search_exact(z[name + "_graph"], z[name + "_codes"], z[name + "_qcodes"],
                                  getattr(O, cdt), z[name + "_raw"], z[name + "_rawq"],
                                  getattr(O, xdt), L, k, rf)
        assert (d.cpu().numpy() == od).all(), (L, k, rf)
    ix.close()


@pytest.mark.parametrize("name,cdt,xdt", WORKLOADS)
def test_frontier_output_and_fast_rerank(gpu, golden_exact, name, cdt, xdt):
    """The whole-frontier output feeds the stand-alone rerank to the fused
    reference-mode result; the FAST form equals its beam re-scored."""
    N, z = gpu, golden_exact
    ix, qd, xb, xq = _exact_index(N, z, name, cdt)
    for L, k, rf in [(20, 10, 1), (20, 10, 3), (40, 10, 100), (32, 5, 2)]:
        fr, _, _ = N.search(ix, qd, L, k, mode=N.SEARCH_REF, frontier=True)
        head, _, _ = N.search(ix, qd, L, k, mode=N.SEARCH_REF)
        assert fr.shape == (len(qd), L)
        assert (fr[:, :k] == head).all()
        a, ad = N.rerank(xb, xq, fr, k, rf)
        b, bd, _ = N.search_rerank(ix, qd, xb, xq, L, k, rf, mode=N.SEARCH_REF)
        assert (a == b).all() and (ad == bd).all(), (L, k, rf)
        beam, _, _ = N.search(ix, qd, L, L, mode=N.SEARCH_FAST)
        c, cd = N.rerank(xb, xq, beam, k, rf)
        f, fd, fst = N.search_rerank(ix, qd, xb, xq, L, k, rf, mode=N.SEARCH_FAST)
        assert (c == f).all() and (cd == fd).all(), (L, k, rf)
        assert (fst[:, 0] > 0).all()
    ix.close()


@pytest.mark.parametrize("name,cdt,xdt", WORKLOADS)
def test_fast_filtering_and_visited_lists(gpu, golden_exact, name, cdt, xdt):
    """FAST filtering mode (rerank_factor 0, beamSearch.h:239-246): each
    query's visited list (NBPG_SEARCH_VISITED: distinct ids, as many as the
    search's visited count, holding the whole final beam) re-scored with the
    raw rows equals the stand-alone rerank of that list, and dist_cmps grows
    by the list's length.  Since the list holds the beam, the i-th exact
    distance of the filtering result is never above the i-th of the beam
    re-scored."""
    import torch
    N, z = gpu, golden_exact
    ix, qd, xb, xq = _exact_index(N, z, name, cdt)
    for L, k in [(16, 10), (20, 10), (40, 5), (128, 100), (256, 10)]:
        cap = N.visited_cap(L)
        vis, _, vst = N.search(ix, qd, L, k, mode=N.SEARCH_FAST, visited=True)
        v = vis.cpu().numpy().view(np.uint32)
        vst = vst.cpu().numpy()
        assert v.shape == (len(qd), cap)
        live = v != 0xFFFFFFFF
        cnt = live.sum(1)
        assert (cnt == np.minimum(vst[:, 0], cap)).all(), (L, k)
        assert all(live[i, :cnt[i]].all() for i in range(len(v)))  # contiguous
        assert all(len(set(v[i, :cnt[i]])) == cnt[i] for i in range(len(v)))
        beam, _, bst = N.search(ix, qd, L, L, mode=N.SEARCH_FAST)
        b = beam.cpu().numpy().view(np.uint32)
        bst = bst.cpu().numpy()
        assert (bst == vst).all(), (L, k)
        for i in range(len(v)):
            bi = b[i][b[i] != 0xFFFFFFFF]
            if cnt[i] < cap:
                assert set(bi) <= set(v[i, :cnt[i]]), (L, k, i)
        f, fd, fst = N.search_rerank(ix, qd, xb, xq, L, k, 0, mode=N.SEARCH_FAST)
        a, ad = N.rerank(xb, xq, vis, k, (cap + k - 1) // k)
        assert torch.equal(f, a) and torch.equal(fd, ad), (L, k)
        fst = fst.cpu().numpy()
        assert (fst[:, 0] == vst[:, 0]).all()
        assert (fst[:, 1] == vst[:, 1] + np.minimum(vst[:, 0], cap)).all(), (L, k)
        c, cd = N.rerank(xb, xq, beam, k, (L + k - 1) // k)
        whole = cnt < cap  # lists that were not cut at the capacity
        assert whole.any()
        assert (fd.cpu().numpy()[whole] <= cd.cpu().numpy()[whole]).all(), (L, k)
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 300, 10, mode=N.SEARCH_FAST, visited=True)
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 20, 10, mode=N.SEARCH_REF, visited=True)
    ix.close()
