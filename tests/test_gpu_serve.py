"""Serving loop: each request's results are the same alone and in a mix of
concurrent requests (bit-identical ids and distances), and equal to a direct
uneven-batch search of the same queries."""
import threading

import numpy as np
import pytest

from nbpg_amd import datagen as D

pytestmark = pytest.mark.gpu

FAST = (2, 1, 11, 0, 0)


@pytest.fixture(scope="module")
def served(gpu):
    N = gpu
    base, q = D.workload_small(30000, 600, dims=64)
    ix = N.Index(N.U8, base.shape[0], base.shape[1], 32)
    ix.set_points(base)
    N.build(ix, 32, 64, alpha=1.2, passes=1, num_split=0)
    ix.host_rows = base  # the oracle checks read the rows the index holds
    rng = np.random.default_rng(9)
    reqs, off = [], 0
    while off < len(q):
        nq = int(min(rng.integers(1, 65), len(q) - off))
        L = int(np.exp(rng.uniform(np.log(16), np.log(256))))
        k = int(rng.choice([1, 10, 100]))
        k = min(k, L)
        reqs.append((q[off:off + nq], L, k))
        off += nq
    yield N, ix, reqs
    ix.close()


def run_alone(N, ix, reqs, **kw):
    srv = N.Server(ix, fast=FAST, **kw)
    out = []
    for rows, L, k in reqs:
        t = srv.submit(rows, L, k)
        ids, d, lat = srv.wait(t)
        assert lat > 0
        out.append((ids, d))
    srv.close()
    return out


def test_same_alone_and_in_mix(served):
    N, ix, reqs = served
    alone = run_alone(N, ix, reqs)
    srv = N.Server(ix, slots=3, max_batch_queries=160, fast=FAST)
    order = np.random.default_rng(1).permutation(len(reqs))
    tickets = {}
    lock = threading.Lock()

    def producer(part):
        for i in part:
            rows, L, k = reqs[i]
            t = srv.submit(rows, L, k)
            with lock:
                tickets[i] = t

    threads = [threading.Thread(target=producer, args=(order[j::4],)) for j in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for i in range(len(reqs)):
        ids, d, _ = srv.wait(tickets[i])
        assert (ids == alone[i][0]).all(), i
        assert (d == alone[i][1]).all(), i
    st = srv.stats()
    assert st["requests"] == len(reqs)
    assert st["queries"] == sum(len(r[0]) for r in reqs)
    assert st["batches"] > 1 and st["mean_batch_queries"] > 20
    assert st["batch_capacity"] == 160
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    srv.close()


def test_matches_direct_uneven_search(served):
    N, ix, reqs = served
    import torch
    alone = run_alone(N, ix, reqs[:12])
    rows = np.concatenate([r[0] for r in reqs[:12]])
    Ls = np.concatenate([[r[1]] * len(r[0]) for r in reqs[:12]]).astype(np.int32)
    ks = np.concatenate([[r[2]] * len(r[0]) for r in reqs[:12]]).astype(np.int32)
    stride = ix.row_stride
    qpad = np.zeros((len(rows), stride), np.uint8)
    qpad[:, :rows.shape[1]] = rows
    ids, d, _ = N.search_uneven(ix, N.to_device(qpad), torch.from_numpy(Ls).cuda(),
                                torch.from_numpy(ks).cuda(), 256, 100, fast=FAST)
    ids = ids.cpu().numpy().view(np.uint32)
    d = d.cpu().numpy()
    off = 0
    for (r, L, k), (ai, ad) in zip(reqs[:12], alone):
        assert (ids[off:off + len(r), :k] == ai).all()
        assert (d[off:off + len(r), :k] == ad).all()
        off += len(r)


def test_memory_admission_and_errors(served):
    N, ix, reqs = served
    stride = ix.row_stride
    per_query = stride + 8 + 8 * 100
    srv = N.Server(ix, slots=2, max_batch_queries=4096, mem_budget=2 * 50 * per_query, fast=FAST)
    assert srv.stats()["batch_capacity"] == 50
    with pytest.raises(N.NbpgError):
        srv.submit(np.zeros((51, 64), np.uint8), 32, 10)  # larger than a batch
    with pytest.raises(N.NbpgError):
        srv.submit(np.zeros((4, 64), np.uint8), 16, 32)  # k > L
    with pytest.raises(N.NbpgError):
        srv.submit(np.zeros((4, 64), np.uint8), 512, 10)  # L > max_L
    rows, L, k = reqs[0]
    t = srv.submit(rows, L, k)
    ids, _, _ = srv.wait(t)
    assert (ids < ix.n).all()
    srv.close()


def _check_against_oracle(base, rows, ids, d, k, gt, gtd):
    """exact distances (the oracle's) per returned id, sorted, distinct;
    returns the hits counted as the reference's recall does"""
    import oracle as O
    from nbpg_amd.autotune import recall_with_ties
    hits = 0.0
    for j in range(len(rows)):
        assert (np.diff(d[j, :k]) >= 0).all()
        assert len(set(ids[j, :k].tolist())) == k
        ref = np.array([O.distance(O.U8, base[i], rows[j]) for i in ids[j, :k]], np.float32)
        assert (ref == d[j, :k]).all(), j
        hits += recall_with_ties(ids[j:j + 1, :k], gt[j:j + 1], gtd[j:j + 1], k) * k
    return hits


def test_uneven_and_served_match_oracle(served):
    """nbpg_search_uneven (mixed L and k in one launch) and the server's
    results against the oracle: each returned id's exact distance, sorted
    distinct ids, and recall@k against the oracle's exact ground truth."""
    N, ix, reqs = served
    import torch
    import oracle as O
    base = ix.host_rows
    sub = reqs[:40]
    rows = np.concatenate([r[0] for r in sub])
    Ls = np.concatenate([[r[1]] * len(r[0]) for r in sub]).astype(np.int32)
    ks = np.concatenate([[r[2]] * len(r[0]) for r in sub]).astype(np.int32)
    gt, gtd = O.compgt(base, rows, 100, O.U8, tie_mode=1)
    stride = ix.row_stride
    qpad = np.zeros((len(rows), stride), np.uint8)
    qpad[:, :rows.shape[1]] = rows
    ids, d, _ = N.search_uneven(ix, N.to_device(qpad), torch.from_numpy(Ls).cuda(),
                                torch.from_numpy(ks).cuda(), 256, 100, fast=FAST)
    ids, d = ids.cpu().numpy().view(np.uint32), d.cpu().numpy()
    hits, want = 0.0, 0
    for j in range(len(rows)):
        k = int(ks[j])
        hits += _check_against_oracle(base, rows[j:j + 1], ids[j:j + 1], d[j:j + 1], k,
                                      gt[j:j + 1], gtd[j:j + 1])
        want += k
    assert hits / want >= 0.95, hits / want
    srv = N.Server(ix, slots=2, max_batch_queries=128, fast=FAST)
    tickets = [srv.submit(r, L, k) for r, L, k in sub]
    off, hits = 0, 0.0
    for (r, L, k), t in zip(sub, tickets):
        si, sd, _ = srv.wait(t)
        hits += _check_against_oracle(base, r, si.view(np.uint32), sd, k,
                                      gt[off:off + len(r)], gtd[off:off + len(r)])
        off += len(r)
    srv.close()
    assert hits / want >= 0.95, hits / want


def test_uneven_L_capped_at_max_L(served):
    """a per-query L above max_L runs as max_L on every kernel shape"""
    N, ix, reqs = served
    import torch
    rows = reqs[0][0][:16]
    stride = ix.row_stride
    qpad = np.zeros((len(rows), stride), np.uint8)
    qpad[:, :rows.shape[1]] = rows
    qd = N.to_device(qpad)
    ks = torch.full((len(rows),), 10, dtype=torch.int32, device="cuda")
    for max_L in (48, 200, 300):
        hi = torch.full((len(rows),), 1000, dtype=torch.int32, device="cuda")
        eq = torch.full((len(rows),), max_L, dtype=torch.int32, device="cuda")
        a = N.search_uneven(ix, qd, hi, ks, max_L, 10, fast=FAST)
        b = N.search_uneven(ix, qd, eq, ks, max_L, 10, fast=FAST)
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), max_L
        assert torch.equal(a[2], b[2]), max_L


def test_served_with_seeds_equals_seeded_searches(served):
    """With 32 k-means start points: each served request equals the seeded
    uneven-batch search and the seeded FAST search of its rows (same L, k),
    and the seeds raise recall@10 over the single entry at the same L."""
    import torch
    import oracle as O
    N, ix, reqs = served
    seeds = N.pick_seeds(ix, 32)
    srv = N.Server(ix, slots=2, max_batch_queries=256, fast=FAST, max_L=256,
                   start_points=seeds)
    for rows, L, k in reqs[:12]:
        t = srv.submit(rows, L, k)
        ids, d, _ = srv.wait(t)
        qpad = np.zeros((len(rows), ix.row_stride), np.uint8)
        qpad[:, :rows.shape[1]] = rows
        qd = N.to_device(qpad)
        Ls = torch.full((len(rows),), L, dtype=torch.int32, device="cuda")
        ks = torch.full((len(rows),), k, dtype=torch.int32, device="cuda")
        ui, ud, _ = N.search_uneven(ix, qd, Ls, ks, L, k, fast=FAST, start_points=seeds)
        fi, fd, _ = N.search(ix, qd, L, k, mode=N.SEARCH_FAST, fast=FAST, start_points=seeds)
        assert (ids == ui.cpu().numpy().view(np.uint32)).all() and (d == ud.cpu().numpy()).all()
        assert (ids == fi.cpu().numpy().view(np.uint32)).all() and (d == fd.cpu().numpy()).all()
    srv.close()
    # recall at a small beam, seeds against node 0
    q = np.concatenate([r for r, _, _ in reqs])[:300]
    qpad = np.zeros((len(q), ix.row_stride), np.uint8)
    qpad[:, :q.shape[1]] = q
    qd = N.to_device(qpad)
    gt, _ = O.compgt(ix.host_rows, q, 10, O.U8, tie_mode=1)
    Ls = torch.full((len(q),), 12, dtype=torch.int32, device="cuda")
    ks = torch.full((len(q),), 10, dtype=torch.int32, device="cuda")
    a, _, _ = N.search_uneven(ix, qd, Ls, ks, 12, 10, fast=FAST, start_points=seeds)
    b, _, _ = N.search_uneven(ix, qd, Ls, ks, 12, 10, fast=FAST)
    ra = O.recall(a.cpu().numpy().view(np.uint32), gt, ix.host_rows, q, O.U8, 10)
    rb = O.recall(b.cpu().numpy().view(np.uint32), gt, ix.host_rows, q, O.U8, 10)
    assert ra >= rb - 0.01, (ra, rb)
    with pytest.raises(N.NbpgError):
        N.Server(ix, fast=FAST, max_L=300, start_points=seeds)


def test_burst_submit_and_wait(served):
    """nbpg_server_submit_many / wait_many: requests queued in bursts return,
    request by request, what each returns alone; bad bursts are refused
    whole, and an unknown ticket stops wait_many with the later tickets still
    waitable one by one."""
    N, ix, reqs = served
    L, k = 48, 10
    rows = np.concatenate([r[0] for r in reqs])
    sizes = np.array([len(r[0]) for r in reqs], np.uint32)
    alone = run_alone(N, ix, [(r[0], L, k) for r in reqs])
    srv = N.Server(ix, slots=2, max_batch_queries=256, fast=FAST)
    off = 0
    got = []
    for lo in range(0, len(sizes), 7):  # bursts of 7 requests, the last shorter
        sz = sizes[lo:lo + 7]
        m = int(sz.sum())
        t = srv.submit_many(rows[off:off + m], sz, L, k)
        ids, d, lat = srv.wait_many(t)
        assert (lat > 0).all()
        got.append((ids.reshape(m, k), d.reshape(m, k)))
        off += m
    ids = np.concatenate([g[0] for g in got])
    d = np.concatenate([g[1] for g in got])
    assert (ids == np.concatenate([a[0] for a in alone])).all()
    assert (d == np.concatenate([a[1] for a in alone])).all()
    with pytest.raises(N.NbpgError):  # one request over the batch capacity refuses all
        srv.submit_many(rows[:300], np.array([10, 290], np.uint32), L, k)
    with pytest.raises(N.NbpgError):  # an empty request
        srv.submit_many(rows[:10], np.array([10, 0], np.uint32), L, k)
    t = srv.submit_many(rows[:30], np.array([10, 20], np.uint32), L, k)
    srv._shape[12345678] = (1, k)
    with pytest.raises(N.NbpgError):
        srv.wait_many(np.array([12345678, t[0], t[1]], np.uint64))
    for tk, want in zip(t.tolist(), (alone[0], None)):
        srv._shape[tk] = (10 if want is not None else 20, k)
        ids1, _, _ = srv.wait(tk)
        if want is not None:
            assert (ids1 == np.concatenate([a[0] for a in alone])[:10]).all()
    srv.close()
