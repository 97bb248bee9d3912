"""The device path against the reference's recorded runs of the branches the
headline configs leave off: RobustPrune's angle test, cosine-threshold and
top_degree builds, searches with limit < 2L, degree_limit < R and L up to
256 (reference_branches.npz; the oracle is pinned to the same fixtures in
test_oracle_branches.py)."""
import numpy as np
import pytest

from test_oracle_branches import BUILDS, live

pytestmark = pytest.mark.gpu


def _index(N, base, dtype, R, rows=None):
    ix = N.Index(dtype, base.shape[0], base.shape[1], R)
    ix.set_points(base)
    if rows is not None:
        ix.set_graph(rows)
    return ix


def test_prune_angle_matches_reference(gpu, golden_branches):
    N = gpu
    g = golden_branches
    base = g["base"]
    R = int(g["prune_R"][0])
    ix = _index(N, base, N.U8, R)
    off = g["prune_offsets"].astype(np.int64)
    cand = N.to_device(N.pack_cands(g["prune_ids"], g["prune_dists"]))
    for si, (alpha, cosa) in enumerate(g["prune_settings"]):
        gi, gd, gc = N.robust_prune_batch(ix, N.to_device(g["prune_nodes"]), cand,
                                          N.to_device(off), R, alpha=float(alpha),
                                          cos_angle=float(cosa))
        gi, gd, gc = gi.cpu().numpy().view(np.uint32), gd.cpu().numpy(), gc.cpu().numpy()
        cnt = g["prune_cnt_%d" % si]
        assert (gc == cnt).all(), si
        for t in range(len(cnt)):
            assert (gi[t, :cnt[t]] == g["prune_out_%d" % si][t, :cnt[t]]).all(), (si, t)
            assert (gd[t, :cnt[t]] == g["prune_outd_%d" % si][t, :cnt[t]]).all(), (si, t)
    ix.close()


@pytest.mark.parametrize("name", sorted(BUILDS))
def test_build_branches_match_reference(gpu, golden_branches, name):
    N = gpu
    g = golden_branches
    dt, key, kw = BUILDS[name]
    kw = dict(kw)
    R = kw.pop("R")
    L = kw.pop("L")
    ix = _index(N, g[key], dt, R)
    N.build(ix, R, L, **kw)
    assert (live(ix.get_graph()) == live(g["graph_" + name])).all(), name
    ix.close()


def test_search_branches_match_reference(gpu, golden_branches):
    N = gpu
    g = golden_branches
    base, q = g["base"], g["queries"]
    ix = _index(N, base, N.U8, 24, g["graph_plain"])
    qpad = np.zeros((len(q), ix.row_stride), np.uint8)
    qpad[:, :q.shape[1]] = q
    qd = N.to_device(qpad)
    for ci, (L, k, lim, dl) in enumerate(g["search_cases"]):
        ids, _, st = N.search(ix, qd, int(L), int(k), mode=N.SEARCH_REF, limit=int(lim),
                              degree_limit=int(dl))
        assert (ids.cpu().numpy().view(np.uint32) == g["search_ids_%d" % ci]).all(), ci
        assert (st.cpu().numpy().view(np.uint32) == g["search_stats_%d" % ci]).all(), ci
    ix.close()
