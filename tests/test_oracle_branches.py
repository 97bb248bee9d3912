"""The oracle's branches that only a recorded run of the reference pins:
RobustPrune's angle test, cosine-threshold and top_degree builds, searches
with limit < 2L, degree_limit < R and wide beams (reference_branches.npz,
written by tests/golden/make_golden_branches.py)."""
import numpy as np

import oracle as O

# Vamana flag sets of the recorded graphs, as build_graph keywords
BUILDS = {
    "cos1": (O.U8, "base", dict(R=24, L=48, alpha=1.2, passes=1, num_split=0,
                                max_cos_angle=0.5)),
    "cos2": (O.U8, "base", dict(R=24, L=40, VL=60, alpha=1.2, min_alpha=1.05, passes=2,
                                num_split=1, dynamic_prune=True, min_cos_angle=0.2,
                                max_cos_angle=0.6)),
    "top8": (O.U8, "base", dict(R=24, L=48, alpha=1.2, passes=2, num_split=0, top_degree=8)),
    "top8u4": (O.U4X2, "base_u4", dict(R=16, L=32, alpha=1.2, passes=2, num_split=0,
                                       top_degree=8)),
    "plain": (O.U8, "base", dict(R=24, L=48, alpha=1.2, passes=1, num_split=0)),
}


def live(rows):
    out = rows.copy()
    R = rows.shape[1] - 1
    mask = np.arange(R)[None, :] >= rows[:, :1]
    out[:, 1:][mask] = 0
    return out


def test_prune_angle_golden(golden_branches):
    g = golden_branches
    base, off = g["base"], g["prune_offsets"]
    R = int(g["prune_R"][0])
    for si, (alpha, cosa) in enumerate(g["prune_settings"]):
        cnt = g["prune_cnt_%d" % si]
        for t, p in enumerate(g["prune_nodes"]):
            b, e = int(off[t]), int(off[t + 1])
            ids, d = O.prune(base, O.U8, int(p), g["prune_ids"][b:e], g["prune_dists"][b:e], R,
                             float(alpha), float(cosa))
            assert len(ids) == cnt[t], (si, t)
            assert (ids == g["prune_out_%d" % si][t, :cnt[t]]).all(), (si, t)
            assert (d == g["prune_outd_%d" % si][t, :cnt[t]]).all(), (si, t)
    # the branch is live: with the angle test off the same lists keep more
    no_angle = sum(len(O.prune(base, O.U8, int(p), g["prune_ids"][int(off[t]):int(off[t + 1])],
                               g["prune_dists"][int(off[t]):int(off[t + 1])], R, 1.2)[0])
                   for t, p in enumerate(g["prune_nodes"]))
    assert g["prune_cnt_1"].sum() < no_angle


def test_builds_golden(golden_branches):
    g = golden_branches
    for name, (dt, key, kw) in BUILDS.items():
        rows, _ = O.build_graph(g[key], dt, **kw)
        assert (live(rows) == live(g["graph_" + name])).all(), name
    # the cosine threshold changes the graph (the branch is live)
    assert not (live(g["graph_cos1"]) == live(g["graph_plain"])).all()


def test_search_limit_degree_limit_wide_beams_golden(golden_branches):
    g = golden_branches
    for ci, (L, k, lim, dl) in enumerate(g["search_cases"]):
        ids, _, st = O.search(g["graph_plain"], g["base"], g["queries"], O.U8, int(L), int(k),
                              limit=int(lim), degree_limit=int(dl))
        assert (ids == g["search_ids_%d" % ci]).all(), ci
        assert (st == g["search_stats_%d" % ci]).all(), ci
