"""The CPU oracle against the reference's recorded outputs and closed forms."""
import os

import numpy as np
import pytest

import oracle as O


def live(rows):
    """(deg, sorted-by-slot live entries) view that ignores stale slots."""
    out = rows.copy()
    R = rows.shape[1] - 1
    mask = np.arange(R)[None, :] >= rows[:, :1]
    out[:, 1:][mask] = 0
    return out


def test_distances_golden(golden):
    a, b = golden["dist_a"], golden["dist_b"]
    for t in range(4):
        got = np.array([O.distance(t, a[i], b[i]) for i in range(len(a))], np.float32)
        assert (got == golden["dist_out"][t]).all()


def test_distances_closed_form():
    z = np.zeros(128, np.uint8)
    f = np.full(128, 255, np.uint8)
    assert O.distance(O.U8, z, f) == 128 * 255 * 255
    assert O.distance(O.I8, z.view(np.int8), f.view(np.int8)) == 128  # 0 vs -1
    # packed 4-bit: a nibble of 0xF vs 0 differs by 240 in the reference's units
    assert O.distance(O.U4X2, np.array([0x0F], np.uint8), np.array([0], np.uint8)) == 240 * 240
    assert O.distance(O.I4X2, np.array([0xFF], np.uint8), np.array([0], np.uint8)) == 2 * 240 * 240
    rng = np.random.default_rng(0)
    x = rng.integers(0, 256, (50, 96), dtype=np.uint8)
    y = rng.integers(0, 256, (50, 96), dtype=np.uint8)
    for i in range(50):
        ref = float(((x[i].astype(np.int64) - y[i]) ** 2).sum())
        assert O.distance(O.U8, x[i], y[i]) == ref
        xs, ys = x[i].view(np.int8).astype(np.int64), y[i].view(np.int8).astype(np.int64)
        assert O.distance(O.I8, x[i].view(np.int8), y[i].view(np.int8)) == float(((xs - ys) ** 2).sum())
        hx, lx = x[i] >> 4, x[i] & 15
        hy, ly = y[i] >> 4, y[i] & 15
        nib = ((hx.astype(np.int64) - hy) ** 2 + (lx.astype(np.int64) - ly) ** 2).sum() * 256
        assert O.distance(O.U4X2, x[i], y[i]) == float(nib)


def test_hash_and_filter_sizes(golden):
    for x, v in zip(golden["hash_in"], golden["hash_out"]):
        assert O.hash64_2(int(x)) == int(v)
    for L, h in zip(golden["filter_L"], golden["filter_size"]):
        assert O.hash_filter_size(int(L)) == int(h)
    # closed form: 2^max(10, ceil(log2 L^2) - 2)
    assert O.hash_filter_size(64) == 1024 and O.hash_filter_size(128) == 4096
    assert O.hash_filter_size(256) == 16384 and O.hash_filter_size(75) == 2048


def test_shuffle_matches_libstdcxx(golden):
    for key in golden:
        if key.startswith("shuffle_sum_"):
            n, seed = map(int, key[len("shuffle_sum_"):].split("_"))
            a = O.shuffle(n, seed)
            assert (a[:2000] == golden["shuffle_%d_%d" % (n, seed)]).all()
            s = int((np.arange(n, dtype=np.uint64) * a.astype(np.uint64)).sum())
            assert s == int(golden[key][0])
            assert sorted(a.tolist()) == list(range(n))


def test_compress_golden_and_closed_form(golden):
    y = golden["compress_y"]
    for sg in (0, 1):
        for ch in (1, 2):
            assert (O.compress(y, sg, ch) == golden["compress_%d_%d" % (sg, ch)]).all()
    # closed form: v=-1 -> 0, v=1 -> 255 (uint, chunk 1); int subtracts 128
    v = np.array([[-1.0, 1.0, 0.0, -0.5]], np.float32)
    u = O.compress(v, 0, 1)[0]
    assert list(u) == [0, 255, 127, 63]
    s = O.compress(v, 1, 1)[0].view(np.int8)
    assert list(s) == [-128, 127, -1, -65]
    # chunk 2 packs dim 2i+1 high, dim 2i low
    p = O.compress(np.array([[1.0, -1.0]], np.float32), 0, 2)[0]
    assert p[0] == 0x0F


def test_compgt_reference_rule_golden(golden):
    ids, d = O.compgt(golden["base"], golden["queries"], 50, O.U8, tie_mode=0)
    assert (ids == golden["gt_ids"]).all()
    assert (d == golden["gt_dists"]).all()


def stale_rule_cases():
    """Hand-worked traces of CompGT.cpp:63-97 for a zero query and rows whose
    squared norms are the listed distances.
    [9, 2, 3, 1, 5], k=3: fill -> max 9 in slot 0; row 3 (d=1) replaces slot
    0, the rescan starts from the pre-write 9 so the max stays 9 at slot 0;
    row 4 (d=5) then evicts slot 0 (the 1), rescan from 1 finds 3 at slot 2.
    Slots {5, 2, 3} -> ids [1, 2, 4], although row 3 is the nearest.
    [5, 5, 3], k=2: row 2 replaces slot 0 (first max), ids [2, 1]; the exact
    (dist, id) rule would keep [2, 0]."""
    rows = {9: [3, 0, 0, 0], 2: [1, 1, 0, 0], 3: [1, 1, 1, 0], 1: [1, 0, 0, 0],
            5: [2, 1, 0, 0]}
    return [
        (np.array([rows[d] for d in [9, 2, 3, 1, 5]], np.uint8), 3, [1, 2, 4], [2, 3, 5]),
        (np.array([rows[d] for d in [5, 5, 3]], np.uint8), 2, [2, 1], [3, 5]),
    ]


def test_compgt_reference_rule_hand_traces():
    for base, k, ids_want, d_want in stale_rule_cases():
        q = np.zeros((1, 4), np.uint8)
        ids, d = O.compgt(base, q, k, O.U8, tie_mode=0)
        assert ids[0].tolist() == ids_want
        assert d[0].tolist() == d_want


def test_compgt_lexicographic_brute_force():
    rng = np.random.default_rng(3)
    base = rng.integers(0, 4, (300, 16), dtype=np.uint8)  # many exact ties
    q = rng.integers(0, 4, (20, 16), dtype=np.uint8)
    ids, d = O.compgt(base, q, 25, O.U8, tie_mode=1)
    for j in range(20):
        dist = ((base.astype(np.int64) - q[j]) ** 2).sum(1)
        order = np.lexsort((np.arange(300), dist))[:25]
        assert (ids[j] == order).all()
        assert (d[j] == dist[order]).all()


def test_compgt_reference_rule_differs_only_by_stale_max():
    """The reference's eviction keeps the multiset of the k smallest only when
    the max never sits in slot 0; check the restatement is at least a subset
    of the prefix and sorted."""
    rng = np.random.default_rng(5)
    base = rng.integers(0, 256, (500, 8), dtype=np.uint8)
    q = rng.integers(0, 256, (10, 8), dtype=np.uint8)
    ids, d = O.compgt(base, q, 10, O.U8, tie_mode=0)
    for j in range(10):
        assert len(set(ids[j].tolist())) == 10
        assert (np.diff(d[j]) >= 0).all()


def test_search_golden(golden):
    g1 = golden["graph1"]
    for i, L in enumerate(golden["search_L"]):
        ids, d, st = O.search(g1, golden["base"], golden["queries"], O.U8, int(L), 10)
        assert (ids == golden["search_ids"][i]).all()
        assert (st == golden["search_stats"][i]).all()


def test_build_golden(golden):
    base = golden["base"]
    g1, _ = O.build_graph(base, O.U8, 24, 48, alpha=1.2, passes=1, num_split=0)
    assert (live(g1) == live(golden["graph1"])).all()
    g2, _ = O.build_graph(base, O.U8, 24, 40, VL=60, alpha=1.1, min_alpha=1.05, passes=2,
                          num_split=1, dynamic_prune=True)
    assert (live(g2) == live(golden["graph2"])).all()


def test_prune_golden(golden):
    base = golden["base"]
    off = golden["prune_offsets"]
    for t, p in enumerate(golden["prune_nodes"]):
        b, e = int(off[t]), int(off[t + 1])
        ids, d = O.prune(base, O.U8, int(p), golden["prune_ids"][b:e], golden["prune_dists"][b:e],
                         16, float(golden["prune_alpha"][t]))
        m = int(golden["prune_cnt"][t])
        assert len(ids) == m
        assert (ids == golden["prune_out"][t, :m]).all()
        assert (d == golden["prune_outd"][t, :m]).all()


def test_prune_hand_worked_chain():
    """Points on a line: 0, 1, 2, 3 (x10).  From p=0 with alpha 1, candidate 1
    occludes 2 and 3 (d(1,2) <= d(0,2)); with alpha 0 nothing is pruned."""
    base = np.zeros((4, 4), np.uint8)
    base[:, 0] = [0, 10, 20, 30]
    ids = np.array([3, 2, 1], np.uint32)
    d = np.array([900, 400, 100], np.float32)
    kept, kd = O.prune(base, O.U8, 0, ids, d, 8, 1.0)
    assert kept.tolist() == [1] and kd.tolist() == [100.0]
    kept, _ = O.prune(base, O.U8, 0, ids, d, 2, 0.0)
    assert kept.tolist() == [1, 2]  # alpha <= 1e-5: sorted, truncated to R
    # alpha 5: 5 * d(1,2) = 500 > 400 keeps 2, 5 * d(1,3) = 2000 > 900 keeps 3
    # past p* = 1, but then p* = 2 prunes 3 since 5 * d(2,3) = 500 <= 900
    kept, _ = O.prune(base, O.U8, 0, ids, d, 8, 5.0)
    assert kept.tolist() == [1, 2]
    kept, _ = O.prune(base, O.U8, 0, ids, d, 8, 10.0)
    assert kept.tolist() == [1, 2, 3]


def test_recall_with_ties():
    base = np.array([[0], [1], [1], [5]], np.uint8)
    q = np.array([[0]], np.uint8)
    gt = np.array([[0, 1, 2, 3]], np.uint32)
    # k = 2: gt[1] has dist 1, gt[2] ties it -> 3 admissible ids
    assert O.recall(np.array([[0, 2]], np.uint32), gt, base, q, O.U8, 2) == 1.0
    assert O.recall(np.array([[0, 3]], np.uint32), gt, base, q, O.U8, 2) == 0.5


def test_pca_compress_reference_golden():
    """matmul + compress == the reference's PCACompress binary, .fbin and all
    four code types, with and without a bias in the .mat."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_pca.npz"))
    for t in "fg":
        y = O.matmul(g["x_" + t], g["W_" + t], g["b_" + t])
        assert (y == g["out_%s_fbin" % t]).all()
        for sg, ch, e in [(0, 1, "u8bin"), (1, 1, "i8bin"), (0, 2, "u4bin"), (1, 2, "i4bin")]:
            assert (O.compress(y, sg, ch) == g["out_%s_%s" % (t, e)]).all(), (t, e)


def test_mat_file_roundtrip(tmp_path):
    from nbpg_amd import pca as P
    W = np.arange(12, dtype=np.float32).reshape(3, 4)
    P.save_mat(str(tmp_path / "a.mat"), W, np.array([1, 2, 3], np.float32))
    W2, b2 = P.load_mat(str(tmp_path / "a.mat"))
    assert (W2 == W).all() and b2.tolist() == [1, 2, 3]
    raw = (tmp_path / "a.mat").read_bytes()
    assert len(raw) == 9 + 4 * 12 + 4 * 3 and raw[:8] == np.array([3, 4], np.uint32).tobytes()
    P.save_mat(str(tmp_path / "b.mat"), W)
    W3, b3 = P.load_mat(str(tmp_path / "b.mat"))
    assert (W3 == W).all() and (b3 == 0).all()


def test_matmul_against_float64():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((40, 96)).astype(np.float32)
    W = rng.standard_normal((16, 96)).astype(np.float32)
    b = rng.standard_normal(16).astype(np.float32)
    y = O.matmul(x, W, b)
    ref = x.astype(np.float64) @ W.T.astype(np.float64) + b
    assert np.abs(y - ref).max() < 1e-4
    # exact sequential fp32 order on a tiny case
    s = np.float32(0)
    for i in range(96):
        s = np.float32(s + np.float32(W[0, i] * x[0, i]))
    assert y[0, 0] == np.float32(b[0] + s)


@pytest.mark.parametrize("name,cdt,xdt", [("f32", O.U8, O.F32), ("u8", O.U4X2, O.U8)])
def test_search_exact_golden(golden_exact, name, cdt, xdt):
    """Rerank (rerank_factor > 0) and filtering (0) searches over quantized
    rows equal the reference's qsearchAll: ids and {visited, dist_cmps}; the
    distances are the raw rows' own (closed form)."""
    z = golden_exact
    xb, xq = z[name + "_raw"], z[name + "_rawq"]
    for i, (L, k, rf) in enumerate(z["settings"]):
        ids, d, st = O.search_exact(z[name + "_graph"], z[name + "_codes"], z[name + "_qcodes"],
                                    cdt, xb, xq, xdt, int(L), int(k), int(rf))
        assert (ids == z[name + "_ids"][i][:, :k]).all(), (L, k, rf)
        assert (st == z[name + "_stats"][i]).all(), (L, k, rf)
        for q in range(0, len(xq), 7):
            for j in range(k):
                if ids[q, j] != 0xFFFFFFFF:
                    assert d[q, j] == O.distance(xdt, xb[ids[q, j]], xq[q])
            live = ids[q] != 0xFFFFFFFF
            assert (np.diff(d[q][live]) >= 0).all()


def test_repair_golden(golden_repair):
    """DFS::repair_isolate_node restated: same rows (including stale slots
    past each degree) and the same isolated count as the reference, on
    graphs that exercise all four repair steps."""
    z = golden_repair
    for g in range(3):
        sp, cnt = (int(v) for v in z["meta%d" % g])
        out, c = O.repair(z["in%d" % g], sp)
        assert c == cnt
        assert (out == z["out%d" % g]).all()


def test_repair_invariants():
    """Closed-form properties of the repair: rows only grow by appended
    entries, every node unreachable from the entry has an in-edge afterwards
    (some reachable row has room here), and new edges point only at
    unreachable nodes."""
    rng = np.random.default_rng(9)
    n, R = 400, 5
    rows = np.zeros((n, R + 1), np.uint32)
    for i in range(n):
        d = int(rng.integers(0, 3))
        rows[i, 0] = d
        rows[i, 1:1 + d] = rng.integers(0, n, d)

    def reach(g):
        seen, stack = {0}, [0]
        while stack:
            v = stack.pop()
            for u in g[v, 1:1 + g[v, 0]]:
                if int(u) not in seen:
                    seen.add(int(u))
                    stack.append(int(u))
        return seen

    before = reach(rows)
    out, cnt = O.repair(rows, 0)
    assert cnt == n - len(before) > 0
    indeg_new = {}
    for v in range(n):
        d0, d1 = int(rows[v, 0]), int(out[v, 0])
        assert d1 >= d0 and (out[v, 1:1 + d0] == rows[v, 1:1 + d0]).all()
        for u in out[v, 1 + d0:1 + d1]:
            indeg_new[int(u)] = indeg_new.get(int(u), 0) + 1
    assert set(indeg_new) <= set(range(n)) - before
    has_in = {int(u) for v in range(n) for u in out[v, 1:1 + out[v, 0]] if int(u) != v}
    assert set(range(n)) - before <= has_in
