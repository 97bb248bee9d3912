"""Write tests/golden/reference_mips.npz: the inner-product metric
(DistanceMetric::MIPS, --use_mips) recorded from the reference's own code in
this container: the Vamana and CompGT binaries and the extern "C" shim
(oracle/_ref, built by oracle/Makefile from the reference tree).

u8 and i8 bases (2000 x 64 mixtures, 40 queries):
  * graphs built by `Vamana --use_mips` (1 pass, alpha 1.2; and 2 passes,
    1 split, dynamic prune), with their search IDs and counters from
    qsearchAll at L = 10, 16, 32, 64
  * `CompGT --use_mips` top-50 (.gt ids and distances)
  * RobustPrune over random candidate lists (alpha 1.2 / 1.0)
  * distances of every dtype, extremes included
  * rerank / filtering searches: f32 rows beside u8 codes

    python tests/golden/make_golden_mips.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from nbpg_amd import datagen as D  # noqa: E402

from make_golden import read_graph, read_gt, write_bin  # noqa: E402

LS = np.array([10, 16, 32, 64], np.uint32)


def main():
    O.build(ref=True)
    R = O.ref()
    assert R is not None, "oracle/_ref not built"
    refdir = os.path.join(ROOT, "oracle", "_ref")
    tmp = tempfile.mkdtemp()
    rng = np.random.default_rng(77)
    out = {"L": LS}
    for name, dt, ext in (("u8", O.U8, "u8bin"), ("i8", O.I8, "i8bin")):
        base = D.mixture_u8(2000, 64, 20, 25.0, seed=41, centers_seed=42)
        queries = D.mixture_u8(40, 64, 20, 25.0, seed=43, centers_seed=42)
        if dt == O.I8:
            base = (base.astype(np.int16) - 128).astype(np.int8).view(np.uint8)
            queries = (queries.astype(np.int16) - 128).astype(np.int8).view(np.uint8)
        bp, qp = os.path.join(tmp, name + "." + ext), os.path.join(tmp, name + "_q." + ext)
        write_bin(bp, base)
        write_bin(qp, queries)
        gtp = os.path.join(tmp, name + ".gt")
        subprocess.run([os.path.join(refdir, "CompGT"), "--base_path", bp, "--query_path", qp,
                        "--gt_path", gtp, "--k", "50", "--use_mips"], check=True,
                       capture_output=True)
        gi, gd = read_gt(gtp)
        graphs = {}
        for gname, args in (("g1", ["--R", "24", "--L", "48", "--max_alpha", "1.2", "--num_split",
                                    "0", "--num_passes", "1"]),
                            ("g2", ["--R", "24", "--L", "40", "--VL", "60", "--num_split", "1",
                                    "--num_passes", "2", "--dynamic_prune", "--min_alpha",
                                    "1.05"])):
            ip = os.path.join(tmp, name + gname)
            subprocess.run([os.path.join(refdir, "Vamana"), "--quant_base_path", bp,
                            "--index_path", ip, "--rebuild", "--use_mips"] + args, check=True,
                           capture_output=True)
            graphs[gname] = read_graph(ip + ".graph")
        g1 = graphs["g1"]
        s_ids = np.zeros((len(LS), len(queries), 10), np.uint32)
        s_st = np.zeros((len(LS), len(queries), 2), np.uint32)
        for i, L in enumerate(LS):
            R.nbpgr_search(dt | O.MIPS, base.ctypes.data, base.shape[0], base.shape[1],
                           g1.ctypes.data, 24, queries.ctypes.data, len(queries), int(L), 10,
                           1.35, base.shape[0], 24, 0, s_ids[i].ctypes.data, s_st[i].ctypes.data)
        # prune over random candidate lists, MIPS distances to the centre
        P = 30
        nodes = rng.integers(0, base.shape[0], P).astype(np.uint32)
        off = np.zeros(P + 1, np.uint64)
        cands = []
        for t in range(P):
            m = int(rng.integers(5, 120))
            ids = rng.integers(0, base.shape[0], m).astype(np.uint32)
            cands.append(ids)
            off[t + 1] = off[t] + m
        cid = np.concatenate(cands)
        cd = np.array([O.distance(dt | O.MIPS, base[j], base[nodes[t]]) for t in range(P)
                       for j in cands[t]], np.float32)
        pR = 16
        alpha = np.array([1.2 if t % 2 else 1.0 for t in range(P)])
        pout = np.full((P, pR), 0xFFFFFFFF, np.uint32)
        pcnt = np.zeros(P, np.uint32)
        for t in range(P):
            b, e = int(off[t]), int(off[t + 1])
            c = np.zeros(e - b, dtype=[("id", np.uint32), ("dist", np.float32)])
            c["id"] = cid[b:e]
            c["dist"] = cd[b:e]
            m = R.nbpgr_prune(dt | O.MIPS, base.ctypes.data, base.shape[0], base.shape[1],
                              int(nodes[t]), c.ctypes.data, e - b, pR, float(alpha[t]), 0.0)
            pcnt[t] = m
            pout[t, :m] = c["id"][:m]
        out.update({name + "_base": base, name + "_queries": queries, name + "_gt_ids": gi,
                    name + "_gt_dists": gd, name + "_graph1": g1, name + "_graph2": graphs["g2"],
                    name + "_search_ids": s_ids, name + "_search_stats": s_st,
                    name + "_prune_nodes": nodes, name + "_prune_offsets": off,
                    name + "_prune_ids": cid, name + "_prune_dists": cd,
                    name + "_prune_alpha": alpha, name + "_prune_out": pout,
                    name + "_prune_cnt": pcnt})
    # distances of every dtype, extremes included
    da = rng.integers(0, 256, (24, 32), dtype=np.uint8)
    db = rng.integers(0, 256, (24, 32), dtype=np.uint8)
    da[0], db[0] = 0, 255
    da[1], db[1] = 255, 255
    da[2], db[2] = 0x80, 0x7F
    out["dist_a"], out["dist_b"] = da, db
    out["dist_out"] = np.array([[R.nbpgr_distance(t | O.MIPS, da[i].ctypes.data, db[i].ctypes.data,
                                                  32) for i in range(len(da))]
                                for t in range(4)], np.float32)
    fa = rng.standard_normal((8, 48)).astype(np.float32)
    fb = rng.standard_normal((8, 48)).astype(np.float32)
    out["fdist_a"], out["fdist_b"] = fa, fb
    out["fdist_out"] = np.array([R.nbpgr_distance(O.F32 | O.MIPS, fa[i].ctypes.data,
                                                  fb[i].ctypes.data, 48) for i in range(8)],
                                np.float32)
    # rerank / filtering with MIPS: f32 rows beside u8 codes, graph over the codes
    xb = (rng.standard_normal((1500, 32)) * np.linspace(3, 0.3, 32)).astype(np.float32)
    xq = (rng.standard_normal((50, 32)) * np.linspace(3, 0.3, 32)).astype(np.float32)
    cb = np.clip(np.rint(xb * 4 + 128), 0, 255).astype(np.uint8)
    cq = np.clip(np.rint(xq * 4 + 128), 0, 255).astype(np.uint8)
    g, _ = O.build_graph(cb, O.U8 | O.MIPS, 16, 32, alpha=1.2, passes=1, num_split=0)
    settings = np.array([(20, 10, 0), (20, 10, 3), (40, 10, 100), (32, 5, 0)], np.uint32)
    e_ids = np.zeros((len(settings), 50, 10), np.uint32)
    e_st = np.zeros((len(settings), 50, 2), np.uint32)
    for i, (L, k, rf) in enumerate(settings):
        oi = np.zeros((50, int(k)), np.uint32)
        os_ = np.zeros((50, 2), np.uint32)
        rc = R.nbpgr_search_exact(O.U8 | O.MIPS, cb.ctypes.data, 1500, 32, g.ctypes.data, 16,
                                  cq.ctypes.data, 50, O.F32, xb.ctypes.data, 32, xq.ctypes.data,
                                  int(L), int(k), 1.35, 1500, 16, 0, int(rf), oi.ctypes.data,
                                  os_.ctypes.data)
        assert rc == 0
        e_ids[i] = 0xFFFFFFFF
        e_ids[i, :, :k] = oi
        e_st[i] = os_
    out.update({"ex_codes": cb, "ex_qcodes": cq, "ex_raw": xb, "ex_rawq": xq, "ex_graph": g,
                "ex_settings": settings, "ex_ids": e_ids, "ex_stats": e_st})
    np.savez_compressed(os.path.join(HERE, "reference_mips.npz"), **out)


if __name__ == "__main__":
    main()
