"""Write tests/golden/reference_branches.npz from the reference's own code:
the branches the other fixtures leave open.

* RobustPrune's angle test (prune.h:200, 218-238) at cos_angle 0.3 and 0.5,
  with alpha 1.2, 1.0 and 0 (angle alone).
* Builds whose prune carries a cosine threshold (getAlphaAndCosAngle,
  nsw.h:297-324): one pass with --max_cos_angle (FixedMinPopular) and a
  dynamic-prune build with --min_cos_angle / --max_cos_angle (PopularPrune
  interpolation, and the default branch's min_cos in the split pass).
* Builds with --top_degree 8 < R (eval_dist_of_new_neighbors, nsw.h:329-355,
  491-517), on u8 and packed u4 rows.
* qsearchAll with limit < 2L (the non-delayed merge, beamSearch.h:121),
  degree_limit < R (beamSearch.h:71) and L = 128, 256.

Runs only in the container that holds the reference tree (oracle/_ref is
built from its sources); the fixtures are data only.

    python tests/golden/make_golden_branches.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from make_golden import read_graph, write_bin  # noqa: E402
from nbpg_amd import datagen as D  # noqa: E402

U4_BASE_ROWS = 2000


def main():
    O.build(ref=True)
    R = O.ref()
    assert R is not None, "oracle/_ref not built"
    refdir = os.path.join(ROOT, "oracle", "_ref")
    tmp = tempfile.mkdtemp()
    rng = np.random.default_rng(77)
    out = {}

    base = D.mixture_u8(2000, 64, 20, 25.0, seed=41, centers_seed=42)
    queries = D.mixture_u8(40, 64, 20, 25.0, seed=43, centers_seed=42)
    b4 = D.mixture_u8(U4_BASE_ROWS, 32, 20, 25.0, seed=44, centers_seed=45)
    out.update(base=base, queries=queries, base_u4=b4)
    write_bin(os.path.join(tmp, "b.u8bin"), base)
    write_bin(os.path.join(tmp, "b4.u4bin"), b4)

    # -- prune with the angle test
    P = 40
    nodes = rng.integers(0, base.shape[0], P).astype(np.uint32)
    offs = np.zeros(P + 1, np.uint64)
    cands = []
    for i in range(P):
        n = int(rng.integers(5, 200))
        ids = rng.integers(0, base.shape[0], n).astype(np.uint32)
        if i % 3 == 0:
            ids[n // 2] = nodes[i]
        cands.append(ids)
        offs[i + 1] = offs[i] + n
    cid = np.concatenate(cands)
    cd = np.array([O.distance(O.U8, base[j], base[nodes[t]]) for t in range(P) for j in cands[t]],
                  np.float32)
    out.update(prune_nodes=nodes, prune_offsets=offs, prune_ids=cid, prune_dists=cd)
    settings = [(1.2, 0.3), (1.2, 0.5), (1.0, 0.5), (0.0, 0.3)]
    out["prune_settings"] = np.array(settings, np.float64)
    pr_R = 24
    out["prune_R"] = np.array([pr_R], np.uint32)
    for si, (alpha, cosa) in enumerate(settings):
        po = np.full((P, pr_R), 0xFFFFFFFF, np.uint32)
        pd = np.zeros((P, pr_R), np.float32)
        pc = np.zeros(P, np.uint32)
        for t in range(P):
            b, e = int(offs[t]), int(offs[t + 1])
            c = np.zeros(e - b, dtype=[("id", np.uint32), ("dist", np.float32)])
            c["id"] = cid[b:e]
            c["dist"] = cd[b:e]
            m = R.nbpgr_prune(0, base.ctypes.data, base.shape[0], base.shape[1], int(nodes[t]),
                              c.ctypes.data, e - b, pr_R, float(alpha), float(cosa))
            pc[t] = m
            po[t, :m] = c["id"][:m]
            pd[t, :m] = c["dist"][:m]
        out["prune_out_%d" % si] = po
        out["prune_outd_%d" % si] = pd
        out["prune_cnt_%d" % si] = pc

    # -- builds (deterministic flag sets: no random start points)
    def vamana(path, args, name):
        subprocess.run([os.path.join(refdir, "Vamana"), "--quant_base_path", path,
                        "--index_path", os.path.join(tmp, name), "--rebuild"] + args, check=True,
                       capture_output=True)
        return read_graph(os.path.join(tmp, name + ".graph"))

    builds = {
        # name: (base file, flags)
        "cos1": ("b.u8bin", ["--R", "24", "--L", "48", "--max_alpha", "1.2", "--num_split", "0",
                             "--num_passes", "1", "--max_cos_angle", "0.5"]),
        "cos2": ("b.u8bin", ["--R", "24", "--L", "40", "--VL", "60", "--num_split", "1",
                             "--num_passes", "2", "--dynamic_prune", "--min_alpha", "1.05",
                             "--max_alpha", "1.2", "--min_cos_angle", "0.2",
                             "--max_cos_angle", "0.6"]),
        "top8": ("b.u8bin", ["--R", "24", "--L", "48", "--max_alpha", "1.2", "--num_split", "0",
                             "--num_passes", "2", "--top_degree", "8"]),
        "top8u4": ("b4.u4bin", ["--R", "16", "--L", "32", "--max_alpha", "1.2", "--num_split",
                                "0", "--num_passes", "2", "--top_degree", "8"]),
    }
    for name, (f, args) in builds.items():
        out["graph_" + name] = vamana(os.path.join(tmp, f), args, name)

    # -- searches over the single-pass cos graph's twin without the angle
    g = vamana(os.path.join(tmp, "b.u8bin"), ["--R", "24", "--L", "48", "--max_alpha", "1.2",
                                              "--num_split", "0", "--num_passes", "1"], "plain")
    out["graph_plain"] = g
    # (L, k, limit, degree_limit)
    cases = [(64, 10, 100, 24), (40, 10, 60, 24), (32, 10, 20, 24), (64, 10, 2000, 12),
             (32, 5, 2000, 6), (48, 10, 90, 16), (128, 10, 2000, 24), (256, 10, 2000, 24),
             (256, 100, 300, 20)]
    out["search_cases"] = np.array(cases, np.uint64)
    for ci, (L, k, lim, dl) in enumerate(cases):
        ids = np.zeros((len(queries), k), np.uint32)
        st = np.zeros((len(queries), 2), np.uint32)
        R.nbpgr_search(0, base.ctypes.data, base.shape[0], base.shape[1], g.ctypes.data, 24,
                       queries.ctypes.data, len(queries), int(L), int(k), 1.35, int(lim), int(dl),
                       0, ids.ctypes.data, st.ctypes.data)
        out["search_ids_%d" % ci] = ids
        out["search_stats_%d" % ci] = st

    path = os.path.join(HERE, "reference_branches.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
