"""The resident graph's .graph I/O and isolated-node repair on the device
(Graph::save / Graph(file), graph.h:169-221, 339-379; DFS::repair_isolate_node,
dfs.h:101-253) against the reference's recorded outputs and the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def live(rows):
    out = rows.copy()
    out[:, 1:][np.arange(rows.shape[1] - 1)[None, :] >= rows[:, :1]] = 0
    return out


def write_graph(path, rows):
    """the reference's .graph layout, written here with numpy"""
    with open(path, "wb") as f:
        np.array([rows.shape[0], rows.shape[1] - 1], np.uint32).tofile(f)
        rows[:, 0].astype(np.uint32).tofile(f)
        mask = np.arange(rows.shape[1] - 1)[None, :] < rows[:, :1]
        rows[:, 1:][mask].astype(np.uint32).tofile(f)


def random_rows(rng, n, R, lo=0):
    rows = np.zeros((n, R + 1), np.uint32)
    rows[:, 0] = rng.integers(lo, R + 1, n)
    rows[:, 1:] = rng.integers(0, n, (n, R))
    return rows


def make_index(N, n, R):
    ix = N.Index(N.U8, n, 32, R)
    ix.set_points(np.zeros((n, 32), np.uint8))
    return ix


def test_repair_matches_reference_golden(gpu, golden_repair):
    N, z = gpu, golden_repair
    for g in range(3):
        sp, cnt = (int(v) for v in z["meta%d" % g])
        rows = z["in%d" % g]
        ix = make_index(N, rows.shape[0], rows.shape[1] - 1)
        ix.set_graph(rows)
        assert ix.repair(sp) == cnt
        assert (live(ix.get_graph()) == live(z["out%d" % g])).all(), g
        ix.close()


def test_repair_matches_oracle_large(gpu):
    """200k nodes, the upper half never pointed at: thousands of isolated
    nodes, long BFS frontiers."""
    N = gpu
    rng = np.random.default_rng(21)
    n, R = 200_000, 12
    rows = np.zeros((n, R + 1), np.uint32)
    rows[:, 0] = rng.integers(0, R + 1, n)
    rows[:, 1:] = rng.integers(0, n // 2, (n, R))
    rows[rng.integers(0, n, 2000), 0] = R  # some full rows
    want, cnt = O.repair(rows, 3)
    ix = make_index(N, n, R)
    ix.set_graph(rows)
    assert ix.repair(3) == cnt > 1000
    assert (live(ix.get_graph()) == live(want)).all()
    # a second repair finds what the first left unreachable and changes only that
    want2, cnt2 = O.repair(want, 3)
    assert ix.repair(3) == cnt2
    assert (live(ix.get_graph()) == live(want2)).all()
    ix.close()


def test_save_and_load_graph(gpu, tmp_path):
    N = gpu
    rng = np.random.default_rng(5)
    n, R = 5000, 24
    rows = random_rows(rng, n, R)
    rows[7, 0] = 0
    ix = make_index(N, n, R)
    ix.set_graph(rows)
    ix.save_graph(tmp_path / "a.graph")
    write_graph(tmp_path / "b.graph", rows)
    assert (tmp_path / "a.graph").read_bytes() == (tmp_path / "b.graph").read_bytes()
    # load into a fresh index, and into one with a larger R
    for R2 in (R, R + 8):
        iy = make_index(N, n, R2)
        iy.load_graph(tmp_path / "b.graph")
        got = live(iy.get_graph())
        assert (got[:, :R + 1] == live(rows)).all() and (got[:, R + 1:] == 0).all()
        iy.close()
    # bad files are errors, not faults
    bad = rows.copy()
    bad[3, 1] = n + 5
    bad[3, 0] = max(bad[3, 0], 1)
    write_graph(tmp_path / "bad_id.graph", bad)
    with pytest.raises(N.NbpgError):
        ix.load_graph(tmp_path / "bad_id.graph")
    write_graph(tmp_path / "small.graph", rows[:100])
    with pytest.raises(N.NbpgError):
        ix.load_graph(tmp_path / "small.graph")
    iz = make_index(N, n, R - 1)
    with pytest.raises(N.NbpgError):
        iz.load_graph(tmp_path / "b.graph")
    iz.close()
    ix.close()


def test_save_and_load_graph_many_chunks(gpu, tmp_path):
    """more ids than one 2^26-id staging step: the id stream is packed and
    unpacked in several node ranges"""
    N = gpu
    rng = np.random.default_rng(6)
    n, R = 1_200_000, 64
    rows = np.empty((n, R + 1), np.uint32)
    rows[:, 0] = R
    rows[::7, 0] = rng.integers(0, R, len(rows[::7]))
    rows[:, 1:] = rng.integers(0, n, (n, R), dtype=np.uint32)
    assert int(rows[:, 0].sum()) > (1 << 26)
    ix = make_index(N, n, R)
    ix.set_graph(rows)
    ix.save_graph(tmp_path / "big.graph")
    write_graph(tmp_path / "ref.graph", rows)
    assert (tmp_path / "big.graph").read_bytes() == (tmp_path / "ref.graph").read_bytes()
    iy = make_index(N, n, R)
    iy.load_graph(tmp_path / "ref.graph")
    assert (live(iy.get_graph()) == live(rows)).all()
    iy.close()
    ix.close()
