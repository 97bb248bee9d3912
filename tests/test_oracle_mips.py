"""The oracle under the inner-product metric (DistanceMetric::MIPS,
point.h:46-156, chosen by --use_mips) against outputs recorded from the
compiled reference, and against closed forms."""
import numpy as np
import pytest

import oracle as O

TYPES = [("u8", O.U8), ("i8", O.I8)]


def live(rows):
    out = rows.copy()
    out[:, 1:][np.arange(rows.shape[1] - 1)[None, :] >= rows[:, :1]] = 0
    return out


def test_mips_distances_golden_and_closed_form(golden_mips):
    z = golden_mips
    for t in range(4):
        got = [O.distance(t | O.MIPS, z["dist_a"][i], z["dist_b"][i]) for i in range(24)]
        assert (np.array(got, np.float32) == z["dist_out"][t]).all(), t
    a, b = z["dist_a"].astype(np.int64), z["dist_b"].astype(np.int64)
    assert (z["dist_out"][0] == -(a * b).sum(1)).all()  # u8: -<a, b>
    sa, sb = a.astype(np.uint8).view(np.int8).astype(np.int64), b.astype(np.uint8).view(np.int8)
    assert (z["dist_out"][1] == -(sa * sb.astype(np.int64)).sum(1)).all()  # i8
    nib = (a & 0xF0) * (b & 0xF0) + ((a << 4) & 0xF0) * ((b << 4) & 0xF0)
    assert (z["dist_out"][2] == -nib.sum(1)).all() and (z["dist_out"][3] == -nib.sum(1)).all()
    fa, fb = z["fdist_a"], z["fdist_b"]
    got = [O.distance(O.F32 | O.MIPS, fa[i], fb[i]) for i in range(8)]
    assert (np.array(got, np.float32) == z["fdist_out"]).all()
    ref = [-float(np.float32(sum(np.float32(x) * np.float32(y) for x, y in zip(fa[i], fb[i]))))
           for i in range(8)]
    assert np.allclose(got, ref, rtol=1e-6)


@pytest.mark.parametrize("name,dt", TYPES)
def test_mips_compgt_search_build_prune_golden(golden_mips, name, dt):
    z = golden_mips
    base, q = z[name + "_base"], z[name + "_queries"]
    ids, d = O.compgt(base, q, 50, dt | O.MIPS, tie_mode=0)
    assert (ids == z[name + "_gt_ids"]).all() and (d == z[name + "_gt_dists"]).all()
    g1 = z[name + "_graph1"]
    for i, L in enumerate(z["L"]):
        sid, _, st = O.search(g1, base, q, dt | O.MIPS, int(L), 10)
        assert (sid == z[name + "_search_ids"][i]).all(), int(L)
        assert (st == z[name + "_search_stats"][i]).all(), int(L)
    b1, _ = O.build_graph(base, dt | O.MIPS, 24, 48, alpha=1.2, passes=1, num_split=0)
    assert (live(b1) == live(g1)).all()
    b2, _ = O.build_graph(base, dt | O.MIPS, 24, 40, VL=60, alpha=1.1, min_alpha=1.05, passes=2,
                          num_split=1, dynamic_prune=True)
    assert (live(b2) == live(z[name + "_graph2"])).all()
    off = z[name + "_prune_offsets"]
    for t, p in enumerate(z[name + "_prune_nodes"]):
        b, e = int(off[t]), int(off[t + 1])
        pid, _ = O.prune(base, dt | O.MIPS, int(p), z[name + "_prune_ids"][b:e],
                         z[name + "_prune_dists"][b:e], 16, float(z[name + "_prune_alpha"][t]))
        m = int(z[name + "_prune_cnt"][t])
        assert len(pid) == m and (pid == z[name + "_prune_out"][t, :m]).all()


def test_mips_exact_search_golden(golden_mips):
    z = golden_mips
    for i, (L, k, rf) in enumerate(z["ex_settings"]):
        ids, d, st = O.search_exact(z["ex_graph"], z["ex_codes"], z["ex_qcodes"], O.U8 | O.MIPS,
                                    z["ex_raw"], z["ex_rawq"], O.F32 | O.MIPS, int(L), int(k),
                                    int(rf))
        assert (ids == z["ex_ids"][i][:, :k]).all(), (L, k, rf)
        assert (st == z["ex_stats"][i]).all(), (L, k, rf)
