"""Write tests/golden/reference_repair.npz: isolated-node repair
(DFS::repair_isolate_node, dfs.h:101-253) recorded from the reference's own
code (oracle/_ref/libnbpg_ref.so, built by oracle/Makefile from the
reference tree in this container).

Inputs are adversarial graphs in the reference's row layout [deg, ids...]:
unreachable components, zero-out-degree nodes, full rows that force the
later repair steps, self loops and an entry whose row is full.

    python tests/golden/make_golden_repair.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402


def graphs():
    rng = np.random.default_rng(404)
    out = []
    # 1: sparse random graph, R = 6, many nodes never pointed at
    n, R = 3000, 6
    rows = np.zeros((n, R + 1), np.uint32)
    for i in range(n):
        d = int(rng.integers(0, R + 1))
        rows[i, 0] = d
        rows[i, 1:1 + d] = rng.choice(n // 2, d, replace=False)  # only the lower half is a target
    out.append((rows, 0))
    # 2: two components, the second unreachable; its rows full or empty
    n, R = 2000, 4
    rows = np.zeros((n, R + 1), np.uint32)
    for i in range(n):
        half = 0 if i < 1000 else 1000
        d = R if i % 3 else int(rng.integers(0, 2))
        rows[i, 0] = d
        rows[i, 1:1 + d] = half + rng.choice(1000, d, replace=False)
    rows[5, 0] = 1
    rows[5, 1] = 5  # a self loop
    out.append((rows, 7))
    # 3: the entry's row is full and most reachable rows are full (step 4)
    n, R = 1500, 3
    rows = np.zeros((n, R + 1), np.uint32)
    for i in range(n):
        if i < 600:
            rows[i, 0] = R
            rows[i, 1:] = rng.choice(600, R, replace=False)
        else:
            d = int(rng.integers(0, 2))
            rows[i, 0] = d
            rows[i, 1:1 + d] = rng.integers(600, n, d)
    rows[599, 0] = 1  # one reachable row with room
    out.append((rows, 0))
    return out


def main():
    O.build(ref=True)
    R = O.ref()
    assert R is not None, "oracle/_ref not built"
    rec = {}
    for g, (rows, sp) in enumerate(graphs()):
        after = rows.copy()
        cnt = R.nbpgr_repair(after.ctypes.data, after.shape[0], after.shape[1] - 1, sp)
        rec["in%d" % g] = rows
        rec["out%d" % g] = after
        rec["meta%d" % g] = np.array([sp, cnt], np.uint64)
    np.savez_compressed(os.path.join(HERE, "reference_repair.npz"), **rec)


if __name__ == "__main__":
    main()
