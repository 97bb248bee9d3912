"""HIP path vs the oracle / the reference's recorded outputs (needs a GPU)."""
import numpy as np
import pytest

import oracle as O
from nbpg_amd import datagen as D

pytestmark = pytest.mark.gpu


def make_index(N, base, dtype, R, rows=None):
    ix = N.Index(dtype, base.shape[0], base.shape[1], R)
    ix.set_points(base)
    if rows is not None:
        ix.set_graph(rows)
    return ix


def dev_rows(N, a, stride):
    """rows padded to `stride` bytes on the device"""
    out = np.zeros((a.shape[0], stride), np.uint8)
    out[:, :a.shape[1]] = a.view(np.uint8)
    return N.to_device(out)


@pytest.fixture(scope="module")
def small(gpu):
    """5000 x 128 u8, 200 queries, R=32 graph built by the oracle."""
    base, q = D.workload_small(5000, 200)
    rows, _ = O.build_graph(base, O.U8, 32, 75, alpha=1.2, passes=1, num_split=0)
    return base, q, rows


def test_l2_batch_exact(gpu):
    N = gpu
    rng = np.random.default_rng(1)
    for dtype, dims in [(N.U8, 128), (N.U8, 96), (N.I8, 96), (N.I8, 64), (N.U4X2, 32),
                        (N.I4X2, 32), (N.U8, 200), (N.U8, 20)]:
        base = rng.integers(0, 256, (1000, dims), dtype=np.uint8)
        base[0] = 0
        base[1] = 255
        base[2] = 0x0F
        q = rng.integers(0, 256, dims, dtype=np.uint8)
        ix = make_index(N, base, dtype, 8)
        ids = np.arange(1000, dtype=np.uint32)
        out = N.l2_batch(ix, N.to_device(q), N.to_device(ids)).cpu().numpy()
        odt = dtype if dtype != N.I4X2 else O.I4X2
        ref = np.array([O.distance(odt, base[i], q) for i in ids], np.float32)
        assert (out == ref).all(), (dtype, dims)
        ix.close()


def test_set_points_strided_and_partial(gpu):
    """Rows uploaded from a wider host stride, in two ranges, land padded
    (the staging + spread path) and give exact distances."""
    N = gpu
    rng = np.random.default_rng(4)
    for dims, hstride in [(20, 24), (100, 100), (96, 112), (20, 32), (50, 64), (120, 128)]:
        wide = rng.integers(0, 256, (3000, hstride), dtype=np.uint8)
        base = wide[:, :dims]
        ix = N.Index(N.U8, 3000, dims, 8)
        for a, b in [(1234, 3000), (0, 1234)]:
            part = np.ascontiguousarray(wide[a:b])
            N._check(N.lib().nbpg_index_set_points(ix.h, part.ctypes.data, hstride, a, b - a),
                     "set_points")
        q = rng.integers(0, 256, dims, dtype=np.uint8)
        ids = np.arange(3000, dtype=np.uint32)
        out = N.l2_batch(ix, N.to_device(q), N.to_device(ids)).cpu().numpy()
        ref = np.array([O.distance(O.U8, base[i], q) for i in ids], np.float32)
        assert (out == ref).all(), dims
        ix.close()


def test_set_points_padded_rows_ignore_padding(gpu):
    """Rows in the reference's padded layout (stride 64 * ceil(bytes / 64),
    point_range.h:51-56, padding left uninitialised there) whose stride
    equals the device row stride: random padding bytes must not reach the
    distances, so l2_batch and a FAST search equal the oracle's."""
    N = gpu
    rng = np.random.default_rng(8)
    for dims, hstride in [(20, 32), (50, 64), (120, 128)]:
        base = D.mixture_u8(3000, dims, 20, 25.0, seed=dims)
        q = D.mixture_u8(40, dims, 20, 25.0, seed=dims + 1)
        wide = rng.integers(0, 256, (3000, hstride), dtype=np.uint8)
        wide[:, :dims] = base
        ix = N.Index(N.U8, 3000, dims, 24)
        N._check(N.lib().nbpg_index_set_points(ix.h, wide.ctypes.data, hstride, 0, 3000),
                 "set_points")
        ids = np.arange(3000, dtype=np.uint32)
        out = N.l2_batch(ix, N.to_device(q[0]), N.to_device(ids)).cpu().numpy()
        ref = np.array([O.distance(O.U8, base[i], q[0]) for i in ids], np.float32)
        assert (out == ref).all(), dims
        rows, _ = O.build_graph(base, O.U8, 24, 48, alpha=1.2)
        ix.set_graph(rows)
        qd = dev_rows(N, q, ix.row_stride)
        fi, fd, _ = N.search(ix, qd, 32, 10, mode=N.SEARCH_FAST)
        fi, fd = fi.cpu().numpy().view(np.uint32), fd.cpu().numpy()
        for j in range(len(q)):
            r = np.array([O.distance(O.U8, base[i], q[j]) for i in fi[j]], np.float32)
            assert (r == fd[j]).all(), (dims, j)
        ri, _, _ = N.search(ix, qd, 32, 10, mode=N.SEARCH_REF)
        oi, _, _ = O.search(rows, base, q, O.U8, 32, 10)
        assert (ri.cpu().numpy().view(np.uint32) == oi).all(), dims
        ix.close()


def test_search_ref_matches_reference_golden(gpu, golden):
    N = gpu
    base, q, g1 = golden["base"], golden["queries"], golden["graph1"]
    ix = make_index(N, base, N.U8, 24, g1)
    qd = dev_rows(N, q, ix.row_stride)
    for i, L in enumerate(golden["search_L"]):
        ids, d, st = N.search(ix, qd, int(L), 10, mode=N.SEARCH_REF)
        assert (ids.cpu().numpy().view(np.uint32) == golden["search_ids"][i]).all(), int(L)
        assert (st.cpu().numpy().view(np.uint32) == golden["search_stats"][i]).all(), int(L)


@pytest.mark.parametrize("L,k,dlim,limit", [(10, 10, 32, None), (16, 10, 32, None),
                                             (32, 10, 32, None), (64, 10, 32, None),
                                             (75, 1, 32, None), (128, 100, 32, None),
                                             (256, 10, 32, None), (64, 10, 20, None),
                                             (64, 10, 32, 100), (40, 0, 32, None)])
def test_search_ref_matches_oracle(gpu, small, L, k, dlim, limit):
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    kk = max(k, 1)
    ids, d, st = N.search(ix, qd, L, k, mode=N.SEARCH_REF, degree_limit=dlim, limit=limit)
    oi, od, ost = O.search(rows, base, q, O.U8, L, k, degree_limit=dlim, limit=limit)
    if k:
        assert (ids.cpu().numpy().view(np.uint32) == oi).all()
        assert (d.cpu().numpy() == od).all()
    assert (st.cpu().numpy().view(np.uint32) == ost).all()
    ix.close()


def test_search_ref_multi_start_and_types(gpu):
    N = gpu
    for dtype, dims, gen in [(N.I8, 64, "i8"), (N.U4X2, 32, "u4"), (N.U8, 96, "u8")]:
        if gen == "i8":
            base = D.mixture_i8(3000, dims, 30, 30.0, seed=5).view(np.uint8)
            q = D.mixture_i8(64, dims, 30, 30.0, seed=6).view(np.uint8)
        else:
            base = D.mixture_u8(3000, dims, 30, 25.0, seed=5)
            q = D.mixture_u8(64, dims, 30, 25.0, seed=6)
        rows, _ = O.build_graph(base, dtype, 24, 48, alpha=1.2)
        ix = make_index(N, base, dtype, 24, rows)
        qd = dev_rows(N, q, ix.row_stride)
        sp = [0, 17, 1234]
        ids, d, st = N.search(ix, qd, 48, 10, mode=N.SEARCH_REF, start_points=sp)
        G = O.Graph(rows.ctypes.data, rows.shape[0], 24)
        oi = np.zeros((len(q), 10), np.uint32)
        od = np.zeros((len(q), 10), np.float32)
        ost = np.zeros((len(q), 2), np.uint32)
        B, Q = O.points(base, dtype), O.points(q, dtype)
        qp = O.QParams(10, 1.35, 48, base.shape[0], 24, 8 * 48)
        spa = np.array(sp, np.uint32)
        assert O.lib().nbpgo_search(O.C.byref(G), O.C.byref(B), O.C.byref(Q), spa.ctypes.data, 3,
                                    O.C.byref(qp), oi.ctypes.data, od.ctypes.data,
                                    ost.ctypes.data, 0) == 0
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), gen
        assert (st.cpu().numpy().view(np.uint32) == ost).all(), gen
        ix.close()


def test_search_ref_rejects_bad_args(gpu, small):
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 8, 10)  # k > L
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 16, 10, entry=base.shape[0])
    ix.close()


def test_search_fast_recall(gpu, small):
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    gt, _ = O.compgt(base, q, 100, O.U8, tie_mode=1)
    for cfg in [(4, 1, 10, 0), (2, 2, 11, 0), (8, 4, 9, 0), (1, 1, 12, 3), (2, 1, 11, 0, 1),
                (4, 2, 10, 0, 1), (2, 4, 10, 512, 1)]:
        ids, d, st = N.search(ix, qd, 64, 10, mode=N.SEARCH_FAST, fast=cfg)
        ids = ids.cpu().numpy().view(np.uint32)
        d = d.cpu().numpy()
        assert (ids < base.shape[0]).all()
        assert (np.diff(d, axis=1) >= 0).all()
        for j in range(len(q)):
            assert len(set(ids[j].tolist())) == 10
            ref = np.array([O.distance(O.U8, base[i], q[j]) for i in ids[j]], np.float32)
            assert (ref == d[j]).all()
        r = O.recall(ids, gt, base, q, O.U8, 10)
        assert r >= 0.95, (cfg, r)
    ix.close()


def test_prune_matches_reference_golden(gpu, golden):
    N = gpu
    base = golden["base"]
    ix = make_index(N, base, N.U8, 24)
    P = len(golden["prune_nodes"])
    cand = N.to_device(N.pack_cands(golden["prune_ids"], golden["prune_dists"]))
    offs = N.to_device(golden["prune_offsets"].astype(np.int64))
    nodes = N.to_device(golden["prune_nodes"])
    alphas = N.to_device(golden["prune_alpha"].astype(np.float32))
    ids, d, cnt = N.robust_prune_batch(ix, nodes, cand, offs, 16, alphas_dev=alphas)
    ids = ids.cpu().numpy().view(np.uint32)
    d = d.cpu().numpy()
    cnt = cnt.cpu().numpy()
    for t in range(P):
        m = int(golden["prune_cnt"][t])
        assert cnt[t] == m
        assert (ids[t, :m] == golden["prune_out"][t, :m]).all()
        assert (d[t, :m] == golden["prune_outd"][t, :m]).all()
    ix.close()


@pytest.mark.parametrize("alpha,cosa,R", [(1.2, 0.0, 64), (1.0, 0.0, 32), (0.0, 0.0, 16),
                                          (1.2, 0.5, 64)])
def test_prune_matches_oracle(gpu, small, alpha, cosa, R):
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32)
    rng = np.random.default_rng(7)
    P = 300
    nodes = rng.integers(0, base.shape[0], P).astype(np.uint32)
    offs = np.zeros(P + 1, np.int64)
    allids, alld = [], []
    for t in range(P):
        n = int(rng.integers(1, 300))
        ids = rng.integers(0, base.shape[0], n).astype(np.uint32)
        if t % 4 == 0:
            ids[n // 2] = nodes[t]
        allids.append(ids)
        alld.append(np.array([O.distance(O.U8, base[i], base[nodes[t]]) for i in ids], np.float32))
        offs[t + 1] = offs[t] + n
    cand = N.to_device(N.pack_cands(np.concatenate(allids), np.concatenate(alld)))
    gi, gd, gc = N.robust_prune_batch(ix, N.to_device(nodes), cand, N.to_device(offs), R,
                                      alpha=alpha, cos_angle=cosa)
    gi, gd, gc = gi.cpu().numpy().view(np.uint32), gd.cpu().numpy(), gc.cpu().numpy()
    for t in range(P):
        oi, od = O.prune(base, O.U8, int(nodes[t]), allids[t], alld[t], R, alpha, cosa)
        assert gc[t] == len(oi), t
        assert (gi[t, :len(oi)] == oi).all(), t
        assert (gd[t, :len(oi)] == od).all(), t
    ix.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_compgt_exact(gpu, mode):
    N = gpu
    base, q = D.workload_small(20000, 100)
    base[10000:10100] = base[:100]  # exact duplicates -> ties broken by id
    q[:5] = base[:5]
    for k in [1, 10, 100]:
        ids, d = N.compgt(N.to_device(base), N.to_device(q), k, N.U8, mode)
        oi, od = O.compgt(base, q, k, O.U8, tie_mode=1)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), k
        assert (d.cpu().numpy() == od).all(), k


def test_compgt_i8_and_4bit(gpu):
    N = gpu
    rng = np.random.default_rng(11)
    for dtype, dims in [(N.I8, 64), (N.U4X2, 32), (N.U8, 96)]:
        base = rng.integers(0, 256, (7000, dims), dtype=np.uint8)
        q = rng.integers(0, 256, (33, dims), dtype=np.uint8)
        ids, d = N.compgt(N.to_device(base), N.to_device(q), 50, dtype, N.GT_PLAIN)
        oi, od = O.compgt(base, q, 50, dtype, tie_mode=1)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), dtype
        assert (d.cpu().numpy() == od).all(), dtype


def test_compgt_mfma_shapes_and_overflow(gpu):
    """MFMA ground truth equals the oracle for i8 / u8 rows of 16..256 bytes,
    ragged n and nq, and when many rows tie so that the candidate lists
    overflow and the plain scan redoes those queries."""
    N = gpu
    rng = np.random.default_rng(5)
    for dtype, dims, n, nq, k in [(N.I8, 64, 5003, 37, 20), (N.U8, 256, 3001, 130, 256),
                                  (N.U8, 16, 20011, 200, 7), (N.I8, 192, 4099, 65, 33)]:
        base = rng.integers(0, 256, (n, dims), dtype=np.uint8)
        q = rng.integers(0, 256, (nq, dims), dtype=np.uint8)
        ids, d = N.compgt(N.to_device(base), N.to_device(q), k, dtype, N.GT_MFMA)
        oi, od = O.compgt(base, q, k, dtype, tie_mode=1)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), (dtype, dims)
        assert (d.cpu().numpy() == od).all(), (dtype, dims)
    # 6000 copies of 3 rows: every query's threshold admits ~2000 rows
    base = np.repeat(rng.integers(0, 256, (3, 128), dtype=np.uint8), 6000, axis=0)
    q = np.concatenate([base[:4], rng.integers(0, 256, (20, 128), dtype=np.uint8)])
    ids, d = N.compgt(N.to_device(base), N.to_device(q), 100, N.U8, N.GT_MFMA)
    oi, od = O.compgt(base, q, 100, N.U8, tie_mode=1)
    assert (ids.cpu().numpy().view(np.uint32) == oi).all()
    assert (d.cpu().numpy() == od).all()
    # base sorted by distance to a point: the strided sample still bounds the
    # threshold from above
    base = rng.integers(0, 256, (400000, 32), dtype=np.uint8)
    base = base[np.argsort(((base.astype(np.int64) - 128) ** 2).sum(1), kind="stable")]
    q = rng.integers(100, 156, (50, 32), dtype=np.uint8)
    ids, d = N.compgt(N.to_device(base), N.to_device(q), 10, N.U8, N.GT_MFMA)
    oi, od = O.compgt(base, q, 10, N.U8, tie_mode=1)
    assert (ids.cpu().numpy().view(np.uint32) == oi).all()
    with pytest.raises(N.NbpgError):
        N.compgt(N.to_device(base[:, :28].copy()), N.to_device(q[:, :28].copy()), 10, N.U8,
                 N.GT_MFMA)


def test_compgt_mfma_bound_too_tight_and_chunks(gpu):
    """Rows near a query placed only where the threshold sample looks (every
    16th row for k = 100): the sample's bound admits fewer than k rows, the
    query goes to the plain redo on the device, and the result is still the
    exact top-k.  More queries than one MFMA chunk (16384) in one call; the
    second call with fewer queries reuses the workspace."""
    N = gpu
    rng = np.random.default_rng(19)
    n, dims = 50000, 64
    base = rng.integers(0, 256, (n, dims), dtype=np.uint8)
    q = rng.integers(0, 256, (40, dims), dtype=np.uint8)
    near = np.clip(q[0].astype(np.int32) + rng.integers(-2, 3, (30, dims)), 0, 255)
    base[16 * np.arange(30)] = near.astype(np.uint8)
    for nqq in (40, 3):
        ids, d = N.compgt(N.to_device(base), N.to_device(q[:nqq]), 100, N.U8, N.GT_MFMA)
        oi, od = O.compgt(base, q[:nqq], 100, O.U8, tie_mode=1)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), nqq
        assert (d.cpu().numpy() == od).all(), nqq
    base = rng.integers(0, 256, (3000, 32), dtype=np.uint8)
    q = rng.integers(0, 256, (20000, 32), dtype=np.uint8)
    for nqq, k in ((20000, 10), (16385, 50), (100, 10)):
        ids, d = N.compgt(N.to_device(base), N.to_device(q[:nqq]), k, N.U8, N.GT_MFMA)
        oi, od = O.compgt(base, q[:nqq], k, O.U8, tie_mode=1)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), (nqq, k)
        assert (d.cpu().numpy() == od).all(), (nqq, k)


@pytest.mark.parametrize("engine", [0, 1])
def test_compgt_reference_rule(gpu, golden, engine):
    """NBPG_GT_REFTIES replays CompGT.cpp:63-97: the compiled reference's own
    output on the golden set, the hand traces of its stale-maximum step, and
    the oracle on tie-heavy and random inputs."""
    N = gpu
    mode = engine | N.GT_REFTIES
    ids, d = N.compgt(N.to_device(golden["base"]), N.to_device(golden["queries"]), 50, N.U8, mode)
    assert (ids.cpu().numpy().view(np.uint32) == golden["gt_ids"]).all()
    assert (d.cpu().numpy() == golden["gt_dists"]).all()
    from test_oracle import stale_rule_cases
    for base, k, ids_want, d_want in stale_rule_cases():
        base16 = np.zeros((len(base), 16), np.uint8)
        base16[:, :4] = base
        ids, d = N.compgt(N.to_device(base16), N.to_device(np.zeros((1, 16), np.uint8)), k, N.U8,
                          mode)
        assert ids.cpu().numpy().view(np.uint32)[0].tolist() == ids_want
        assert d.cpu().numpy()[0].tolist() == d_want
    rng = np.random.default_rng(21)
    cases = [(N.U8, 16, 3000, 70, 25, 4), (N.U8, 128, 20000, 100, 100, 256),
             (N.I8, 64, 9001, 33, 200, 256), (N.U8, 32, 5000, 40, 1, 3)]
    if engine == 0:
        cases += [(N.U8, 12, 4000, 20, 10, 3), (N.U4X2, 32, 6000, 25, 50, 256)]
    for dtype, dims, n, nq, k, hi in cases:
        base = rng.integers(0, hi, (n, dims), dtype=np.uint8)
        q = rng.integers(0, hi, (nq, dims), dtype=np.uint8)
        ids, d = N.compgt(N.to_device(base), N.to_device(q), k, dtype, mode)
        oi, od = O.compgt(base, q, k, dtype, tie_mode=0)
        assert (ids.cpu().numpy().view(np.uint32) == oi).all(), (dtype, dims, k)
        assert (d.cpu().numpy() == od).all(), (dtype, dims, k)


def live(rows):
    out = rows.copy()
    R = rows.shape[1] - 1
    out[:, 1:][np.arange(R)[None, :] >= rows[:, :1]] = 0
    return out


def test_build_matches_reference_golden(gpu, golden):
    """Device build == the reference's own Vamana binary, row for row."""
    N = gpu
    base = golden["base"]
    ix = make_index(N, base, N.U8, 24)
    N.build(ix, 24, 48, alpha=1.2, passes=1, num_split=0)
    assert (live(ix.get_graph()) == live(golden["graph1"])).all()
    ix.close()
    ix = make_index(N, base, N.U8, 24)
    N.build(ix, 24, 40, VL=60, alpha=1.1, min_alpha=1.05, passes=2, num_split=1,
            dynamic_prune=True)
    assert (live(ix.get_graph()) == live(golden["graph2"])).all()
    ix.close()


@pytest.mark.parametrize("dtype,dims,kw", [
    ("u8", 128, dict(R=32, L=64, alpha=1.2, passes=1, num_split=0)),
    ("u8", 128, dict(R=32, L=50, alpha=1.1, min_alpha=1.05, passes=2, num_split=4,
                     dynamic_prune=True, sp_seed=7)),
    ("i8", 96, dict(R=24, L=40, alpha=1.2, passes=1, num_split=0, rand_sp=True)),
    ("u4", 32, dict(R=16, L=32, alpha=1.2, passes=2, num_split=0, top_degree=8)),
])
def test_build_matches_oracle(gpu, dtype, dims, kw):
    N = gpu
    if dtype == "i8":
        base = D.mixture_i8(4000, dims, 40, 30.0, seed=31).view(np.uint8)
        t = N.I8
    else:
        base = D.mixture_u8(4000, dims, 40, 25.0, seed=31)
        t = N.U8 if dtype == "u8" else N.U4X2
    R = kw.pop("R")
    L = kw.pop("L")
    ix = make_index(N, base, t, R)
    iso = N.build(ix, R, L, **kw)
    rows, oiso = O.build_graph(base, t, R, L, **kw)
    assert iso == oiso
    assert (live(ix.get_graph()) == live(rows)).all()
    ix.close()


def test_fast_search_hipgraph_replay(gpu, small):
    """One fast search captured in a hipGraph; every replay (with eager work
    in between) reproduces the eager result."""
    import torch
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    e_ids, e_d, e_st = N.search(ix, qd, 48, 10, mode=N.SEARCH_FAST)
    e_ids, e_d = e_ids.cpu().numpy(), e_d.cpu().numpy()
    out = N.search(ix, qd, 48, 10, mode=N.SEARCH_FAST)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        N.search(ix, qd, 48, 10, mode=N.SEARCH_FAST, out=out)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        N.search(ix, qd, 48, 10, mode=N.SEARCH_FAST, out=out)
    for rep in range(4):
        out[0].fill_(-1)
        out[1].fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        assert (out[0].cpu().numpy() == e_ids).all(), rep
        assert (out[1].cpu().numpy() == e_d).all(), rep
        N.search(ix, qd, 32, 10, mode=N.SEARCH_REF)  # eager work between replays
        torch.cuda.synchronize()
    ix.close()


def test_fast_and_reference_modes_agree_mostly(gpu, small):
    """Both modes are best-first searches over the same graph: at a beam
    width with high recall their top-10 sets nearly coincide."""
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    a, _, _ = N.search(ix, qd, 64, 10, mode=N.SEARCH_FAST)
    b, _, _ = N.search(ix, qd, 64, 10, mode=N.SEARCH_REF)
    a, b = a.cpu().numpy(), b.cpu().numpy()
    agree = np.mean([len(set(x) & set(y)) / 10 for x, y in zip(a, b)])
    assert agree > 0.97, agree
    ix.close()


@pytest.mark.parametrize("L,nseeds", [(16, 64), (48, 20), (130, 64), (10, 2)])
def test_fast_search_seeds(gpu, small, L, nseeds):
    """FAST search from several start nodes: exact sorted distances, the
    seeds' own distances counted, recall at least that of the single entry
    (from 16 seeds),
    the argument checks (repeats, more than 64), and the same search on the
    row-split kernel."""
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    seeds = N.pick_seeds(ix, nseeds)
    assert len(seeds) == len(np.unique(seeds)) and len(seeds) <= nseeds
    gt, gd = O.compgt(base, q, 100, O.U8, tie_mode=1)
    ids, d, st = N.search(ix, qd, L, 10, mode=N.SEARCH_FAST, start_points=seeds)
    ids, d, st = ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), st.cpu().numpy()
    assert (np.diff(d, axis=1) >= 0).all()
    for j in range(0, len(q), 9):
        ref = np.array([O.distance(O.U8, base[i], q[j]) for i in ids[j]], np.float32)
        assert (ref == d[j]).all()
    assert (st[:, 1] >= len(seeds)).all()
    if L <= 16 and len(seeds) >= L:  # the seeds alone fill the beam
        for j in range(len(q)):
            sd = sorted((O.distance(O.U8, base[i], q[j]), i) for i in seeds)
            assert d[j, 0] <= sd[0][0]
    e0, _, _ = N.search(ix, qd, L, 10, mode=N.SEARCH_FAST)
    r_seed = O.recall(ids, gt, base, q, O.U8, 10)
    r_e0 = O.recall(e0.cpu().numpy().view(np.uint32), gt, base, q, O.U8, 10)
    if nseeds >= 16:  # two seeds can start farther away than node 0
        assert r_seed >= r_e0 - 0.01, (r_seed, r_e0)
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, L, 10, mode=N.SEARCH_FAST, start_points=[1, 2, 2])
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, L, 10, mode=N.SEARCH_FAST, start_points=list(range(65)))
    # the row-split kernel from the same seeds: exact sorted distances, the
    # seeds counted, recall on par with the one-lane-per-row kernel
    si, sdist, sst = N.search(ix, qd, L, 10, mode=N.SEARCH_FAST, start_points=seeds,
                              fast=(2, 1, 11, 0, 1))
    si, sdist, sst = si.cpu().numpy().view(np.uint32), sdist.cpu().numpy(), sst.cpu().numpy()
    assert (np.diff(sdist, axis=1) >= 0).all()
    for j in range(0, len(q), 9):
        ref = np.array([O.distance(O.U8, base[i], q[j]) for i in si[j]], np.float32)
        assert (ref == sdist[j]).all()
        assert len(set(si[j].tolist())) == 10
    assert (sst[:, 1] >= len(seeds)).all()
    assert O.recall(si, gt, base, q, O.U8, 10) >= r_seed - 0.01
    ix.close()


@pytest.mark.parametrize("L,cfg", [(300, (2, 1, 10, 0)), (200, (4, 2, 11, 0)), (100, (2, 4, 9, 0)),
                                   (200, (2, 1, 11, 0, 1)), (100, (4, 4, 10, 0, 1))])
def test_search_fast_large_beams(gpu, small, L, cfg):
    """Register beams (L <= 256, 1-4 entries per lane) and the LDS-resident
    beam (L > 256) return exact distances, sorted, and high recall."""
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    gt, _ = O.compgt(base, q, 100, O.U8, tie_mode=1)
    ids, d, st = N.search(ix, qd, L, 50, mode=N.SEARCH_FAST, fast=cfg)
    ids = ids.cpu().numpy().view(np.uint32)
    d = d.cpu().numpy()
    assert (np.diff(d, axis=1) >= 0).all()
    for j in range(0, len(q), 7):
        ref = np.array([O.distance(O.U8, base[i], q[j]) for i in ids[j]], np.float32)
        assert (ref == d[j]).all()
        assert len(set(ids[j].tolist())) == 50
    assert O.recall(ids, gt, base, q, O.U8, 50) >= 0.97
    ix.close()


@pytest.mark.parametrize("cfg", [(2, 1, 10, 0, 0), (4, 2, 11, 0, 0), (2, 1, 10, 0, 1)])
def test_fast_results_independent_of_batch_position(gpu, small, cfg):
    """A FAST query's result depends on the query, the index and the launch
    shape only: permuting the batch, shrinking it or changing the number of
    persistent workgroups (which changes which wave runs it, from its grid
    index or the work counter) returns the same ids, distances and counters."""
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    seeds = N.pick_seeds(ix, 16)
    L = 40
    ids, d, st = N.search(ix, dev_rows(N, q, ix.row_stride), L, 10, mode=N.SEARCH_FAST, fast=cfg,
                          start_points=seeds)
    ids, d, st = (t.cpu().numpy() for t in (ids, d, st))
    perm = np.random.default_rng(3).permutation(len(q))
    for qs, sel, groups in ((q[perm], perm, 0), (q[:37], np.arange(37), 0), (q, np.arange(len(q)), 3)):
        c = (cfg[0], cfg[1], cfg[2], groups, cfg[4])
        i2, d2, s2 = N.search(ix, dev_rows(N, qs, ix.row_stride), L, 10, mode=N.SEARCH_FAST,
                              fast=c, start_points=seeds)
        assert (i2.cpu().numpy() == ids[sel]).all()
        assert (d2.cpu().numpy() == d[sel]).all()
        assert (s2.cpu().numpy() == st[sel]).all()
    ix.close()


@pytest.mark.parametrize("gen", ["i8", "u4", "i4"])
@pytest.mark.parametrize("cfg", [(2, 1, 10, 0, 0), (2, 1, 10, 0, 1), (4, 2, 11, 0, 0)])
def test_search_fast_other_types(gpu, gen, cfg):
    """FAST search over i8 and packed 4-bit rows, both kernels, from seeds:
    exact distances per returned id (the oracle's), sorted, distinct ids,
    recall against the oracle's ground truth."""
    N = gpu
    dt = {"i8": N.I8, "u4": N.U4X2, "i4": N.I4X2}[gen]
    if gen == "i8":
        base = D.mixture_i8(4000, 64, 30, 30.0, seed=15).view(np.uint8)
        q = D.mixture_i8(80, 64, 30, 30.0, seed=16).view(np.uint8)
    else:
        base = D.mixture_u8(4000, 32, 30, 25.0, seed=15)
        q = D.mixture_u8(80, 32, 30, 25.0, seed=16)
    rows, _ = O.build_graph(base, dt, 24, 48, alpha=1.2)
    ix = make_index(N, base, dt, 24, rows)
    qd = dev_rows(N, q, ix.row_stride)
    seeds = N.pick_seeds(ix, 32)
    ids, d, _ = N.search(ix, qd, 48, 10, mode=N.SEARCH_FAST, fast=cfg, start_points=seeds)
    ids, d = ids.cpu().numpy().view(np.uint32), d.cpu().numpy()
    assert (np.diff(d, axis=1) >= 0).all()
    for j in range(len(q)):
        assert len(set(ids[j].tolist())) == 10
        ref = np.array([O.distance(dt, base[i], q[j]) for i in ids[j]], np.float32)
        assert (ref == d[j]).all(), j
    gt, _ = O.compgt(base, q, 100, dt, tie_mode=1)
    assert O.recall(ids, gt, base, q, dt, 10) >= 0.9, (gen, cfg)
    ix.close()


@pytest.mark.parametrize("cfg", [(2, 1, 10, 0, 0), (2, 2, 11, 0, 1), (2, 1, 11, 0, 0)])
def test_fast_edge_counts(gpu, cfg):
    """NBPG_SEARCH_EDGES: the third stats word is the adjacency entries of
    the expanded nodes.  On graphs where every node has the same degree d it
    equals visited * d; ids, distances and the first two words are those of
    the plain call.  L = 300 takes the LDS-beam kernel."""
    N = gpu
    rng = np.random.default_rng(31)
    n, R = 3000, 24
    base = D.mixture_u8(n, 64, 20, 25.0, seed=5)
    q = D.mixture_u8(50, 64, 20, 25.0, seed=6)
    for d in (R, 7):
        rows = np.zeros((n, R + 1), np.uint32)
        rows[:, 0] = d
        for i in range(n):
            nb = rng.choice(n - 1, d, replace=False)
            rows[i, 1:1 + d] = nb + (nb >= i)
        ix = make_index(N, base, N.U8, R, rows)
        qd = dev_rows(N, q, ix.row_stride)
        for L in (32, 300):
            a = N.search(ix, qd, L, 10, mode=N.SEARCH_FAST, fast=cfg)
            b = N.search(ix, qd, L, 10, mode=N.SEARCH_FAST, fast=cfg, edges=True)
            assert (a[0].cpu() == b[0].cpu()).all() and (a[1].cpu() == b[1].cpu()).all()
            sa, sb = a[2].cpu().numpy(), b[2].cpu().numpy()
            assert sb.shape == (len(q), 3)
            assert (sa == sb[:, :2]).all(), (d, L)
            assert (sb[:, 2] == sb[:, 0] * d).all(), (d, L)
        with pytest.raises(N.NbpgError):
            N.search(ix, qd, 32, 10, mode=N.SEARCH_REF, edges=True)
        ix.close()


@pytest.mark.parametrize("mode,L", [("fast", 32), ("ref", 32), ("ref", 256), ("fast", 300)])
def test_search_graph_replay_equals_eager(gpu, small, mode, L):
    """nbpg_search_graph_*: the captured search replayed several times, with
    eager searches and synchronisations between the replays and new query
    rows written into the captured buffer, equals the eager search of the
    same rows every time (ids, distances, counters)."""
    import torch
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    m = N.SEARCH_FAST if mode == "fast" else N.SEARCH_REF
    seeds = N.pick_seeds(ix, 16) if mode == "fast" and L <= 256 else None
    qa = dev_rows(N, q[:100], ix.row_stride)
    qb = dev_rows(N, q[100:200], ix.row_stride)
    buf = qa.clone()
    g = N.SearchGraph(ix, buf, L, 10, mode=m, start_points=seeds)
    for rep in range(4):
        src = qa if rep % 2 == 0 else qb
        buf.copy_(src)
        got = [t.clone() for t in g.replay()]
        torch.cuda.synchronize()
        want = N.search(ix, src, L, 10, mode=m, start_points=seeds)
        torch.cuda.synchronize()
        for a, b in zip(got, want):
            assert torch.equal(a, b), (mode, L, rep)
    g.close()
    ix.close()


def test_ref_search_graph_owns_its_scratch(gpu, small):
    """A reference-mode graph whose per-query state lives outside LDS (L 256:
    a 64 KB filter) keeps that state in scratch of its own: eager searches
    that grow and reallocate the index's workspace between replays (more
    queries, larger L) leave every replay equal to the eager search."""
    import torch
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qa = dev_rows(N, q[:100], ix.row_stride)
    qall = dev_rows(N, q, ix.row_stride)
    buf = qa.clone()
    g = N.SearchGraph(ix, buf, 256, 10, mode=N.SEARCH_REF)
    for rep, L in enumerate((300, 400)):
        N.search(ix, qall, L, 10, mode=N.SEARCH_REF)  # grows the workspace
        torch.cuda.synchronize()
        got = [t.clone() for t in g.replay()]
        torch.cuda.synchronize()
        want = N.search(ix, qa, 256, 10, mode=N.SEARCH_REF)
        torch.cuda.synchronize()
        for a, b in zip(got, want):
            assert torch.equal(a, b), (L, rep)
    g.close()
    ix.close()


def test_uneven_search_graph_replay_equals_eager(gpu, small):
    """nbpg_search_uneven_graph_*: the captured uneven search replayed with new
    query rows, beam widths and k written into the captured buffers between
    replays (eager searches and synchronisations in between) equals the eager
    nbpg_search_uneven of the same inputs every time."""
    import torch
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    seeds = N.pick_seeds(ix, 16)
    rng = np.random.default_rng(23)
    qa = dev_rows(N, q[:100], ix.row_stride)
    qb = dev_rows(N, q[100:200], ix.row_stride)

    def lk(seed):
        r = np.random.default_rng(seed)
        L = np.exp(r.uniform(np.log(16), np.log(256), 100)).round().astype(np.uint32)
        k = np.minimum(r.choice(np.array([1, 10, 100], np.uint32), 100), L)
        return (torch.from_numpy(L.view(np.int32)).cuda(), torch.from_numpy(k.view(np.int32)).cuda())
    qbuf, (Lbuf, kbuf) = qa.clone(), lk(1)
    g = N.UnevenSearchGraph(ix, qbuf, Lbuf, kbuf, 256, 100, start_points=seeds)
    for rep in range(4):
        src = qa if rep % 2 == 0 else qb
        Ls, ks = lk(int(rng.integers(1 << 30)))
        qbuf.copy_(src)
        Lbuf.copy_(Ls)
        kbuf.copy_(ks)
        got = [t.clone() for t in g.replay()]
        torch.cuda.synchronize()
        want = N.search_uneven(ix, src, Ls, ks, 256, 100, start_points=seeds)
        torch.cuda.synchronize()
        for a, b in zip(got, want):
            assert torch.equal(a, b), rep
    g.close()
    ix.close()


@pytest.mark.parametrize("kind", ["fast", "uneven"])
def test_concurrent_searches_on_streams(gpu, small, kind):
    """Searches on one index issued round-robin on three streams, 70 of them
    (more than the index's 64 work counters, so every counter is reused),
    plus two search graphs replayed on a fourth stream in between: every
    launch returns what the same search returns alone."""
    import torch
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    seeds = N.pick_seeds(ix, 16)
    batches = [dev_rows(N, q[i * 50:(i + 1) * 50], ix.row_stride) for i in range(4)]
    Ls = torch.tensor(np.full(50, 24, np.uint32).view(np.int32), device="cuda")
    ks = torch.tensor(np.full(50, 10, np.uint32).view(np.int32), device="cuda")

    def run(b, out=None):
        if kind == "fast":
            return N.search(ix, batches[b], 24, 10, mode=N.SEARCH_FAST, start_points=seeds,
                            out=out)
        return N.search_uneven(ix, batches[b], Ls, ks, 24, 10, out=out)
    want = []
    for b in range(4):
        want.append([t.clone() for t in run(b)])
        torch.cuda.synchronize()
    graphs = [N.SearchGraph(ix, batches[b], 24, 10, mode=N.SEARCH_FAST, start_points=seeds)
              for b in (0, 1)] if kind == "fast" else []
    streams = [torch.cuda.Stream() for _ in range(4)]
    outs = [[torch.empty_like(t) for t in want[i % 4]] for i in range(70)]
    torch.cuda.synchronize()
    for i in range(70):
        with torch.cuda.stream(streams[i % 3]):
            run(i % 4, outs[i])
        if graphs and i % 10 == 5:
            with torch.cuda.stream(streams[3]):
                graphs[(i // 10) % 2].replay()
    torch.cuda.synchronize()
    for i in range(70):
        for a, b in zip(outs[i], want[i % 4]):
            assert torch.equal(a, b), (kind, i)
    for j, g in enumerate(graphs):
        for a, b in zip((g.ids, g.dists, g.stats), want[j]):
            assert torch.equal(a, b), ("graph", j)
        g.close()
    ix.close()


def _nearest_seed(base, q, seeds, dtype, mips):
    """each query's nearest seed by exact integer distance, ties to the lower
    id (seeds ascending): the restatement the device pass must equal"""
    s = np.sort(seeds)
    if dtype == "i8":
        b, qq = base[s].view(np.int8).astype(np.int64), q.view(np.int8).astype(np.int64)
    else:
        b, qq = base[s].astype(np.int64), q.astype(np.int64)
    d = -(qq @ b.T) if mips else ((qq[:, None, :] - b[None, :, :]) ** 2).sum(-1)
    return s[np.argmin(d, axis=1)].astype(np.uint32)


@pytest.mark.parametrize("dtype,mips", [("u8", False), ("i8", False), ("i8", True), ("u8", True)])
def test_nearest_seed_starts(gpu, small, dtype, mips):
    """NBPG_SEARCH_SEED_NEAREST: every query starts from its nearest seed row
    (exact distances, ties to the lower id, 300 seeds: a padded last tile),
    and returns exactly what the same FAST search returns when started from
    that seed alone (ids, distances, counters), in both metrics."""
    import torch
    N = gpu
    base, q, rows = small
    rng = np.random.default_rng(41)
    if dtype == "i8":
        base = base[:, :64] ^ 0x80  # 64-byte signed rows around zero
        q = q[:, :64] ^ 0x80
        dt, R = N.I8, 32
    else:
        dt, R = N.U8, 32
    ix = make_index(N, base, dt, R, rows)
    if mips:
        ix.set_metric(N.METRIC_MIPS)
    seeds = rng.choice(base.shape[0], 300, replace=False).astype(np.uint32)
    ix.set_seeds(seeds)
    qd = dev_rows(N, q, ix.row_stride)
    mode = N.SEARCH_FAST | N.SEARCH_SEED_NEAREST
    for L in (16, 64):
        ids, d, st = N.search(ix, qd, L, 10, mode=mode)
        ids, d, st = ids.cpu().numpy(), d.cpu().numpy(), st.cpu().numpy()
        want_start = _nearest_seed(base, q, seeds, dtype, mips)
        for s in np.unique(want_start):
            sel = np.nonzero(want_start == s)[0]
            qs = qd[torch.from_numpy(sel).cuda()]
            wi, wd, ws = N.search(ix, qs, L, 10, mode=N.SEARCH_FAST, start_points=[int(s)])
            assert (wi.cpu().numpy() == ids[sel]).all(), (dtype, mips, L, int(s))
            assert (wd.cpu().numpy() == d[sel]).all(), (dtype, mips, L, int(s))
            assert (ws.cpu().numpy() == st[sel]).all(), (dtype, mips, L, int(s))
    ix.close()


def test_nearest_seed_starts_errors(gpu, small):
    """refused: duplicate or out-of-range seeds, too many, searches without
    seeds (also after new rows drop them), with a visited list, k = 0 or a
    4-bit index"""
    N = gpu
    base, q, rows = small
    ix = make_index(N, base, N.U8, 32, rows)
    qd = dev_rows(N, q, ix.row_stride)
    mode = N.SEARCH_FAST | N.SEARCH_SEED_NEAREST
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 16, 10, mode=mode)  # no seeds yet
    with pytest.raises(N.NbpgError):
        ix.set_seeds([3, 5, 3])
    with pytest.raises(N.NbpgError):
        ix.set_seeds([base.shape[0]])
    with pytest.raises(N.NbpgError):
        ix.set_seeds(np.arange(N.MAX_SEEDS + 1))
    ix.set_seeds([7, 1, 4000])
    ids, _, _ = N.search(ix, qd, 16, 10, mode=mode)
    assert (ids.cpu().numpy().view(np.uint32) < base.shape[0]).all()
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 16, 10, mode=mode, visited=True)
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 16, 0, mode=mode)
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 16, 10, mode=N.SEARCH_REF | N.SEARCH_SEED_NEAREST)
    ix.set_points(base[:10])  # new rows drop the seeds
    with pytest.raises(N.NbpgError):
        N.search(ix, qd, 16, 10, mode=mode)
    ix.set_seeds([])
    ix.close()
    rng = np.random.default_rng(3)
    b4 = rng.integers(0, 256, (500, 32), dtype=np.uint8)
    ix4 = make_index(N, b4, N.U4X2, 8)
    with pytest.raises(N.NbpgError):
        ix4.set_seeds([1, 2])
    ix4.close()
