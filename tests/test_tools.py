"""The command-line tools: flag handling on the host; on the device, their
files equal the reference binaries' own outputs (golden fixtures)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
BIN = os.path.join(ROOT, "nbpg_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.skip("tools not built")
    return p


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=120, **kw)


def write_bin(path, a):
    with open(path, "wb") as f:
        np.array([a.shape[0], a.shape[1]], np.uint32).tofile(f)
        np.ascontiguousarray(a).tofile(f)


def read_graph(path):
    raw = np.fromfile(path, dtype=np.uint32)
    n, R = int(raw[0]), int(raw[1])
    deg = raw[2:2 + n]
    rows = np.zeros((n, R + 1), np.uint32)
    rows[:, 0] = deg
    off = 2 + n
    for i in range(n):
        rows[i, 1:1 + deg[i]] = raw[off:off + deg[i]]
        off += deg[i]
    assert off == len(raw)
    return rows


def live(rows):
    out = rows.copy()
    out[:, 1:][np.arange(rows.shape[1] - 1)[None, :] >= rows[:, :1]] = 0
    return out


@pytest.mark.parametrize("name", ["Vamana", "CompGT", "PCACompress"])
def test_unknown_and_missing_flags_are_errors(name):
    r = run([tool(name), "--no_such_flag", "3"])
    assert r.returncode == 1 and "unknown option --no_such_flag" in r.stderr
    r = run([tool(name)])
    assert r.returncode == 1 and "required" in r.stderr


def test_refused_options():
    r = run([tool("CompGT"), "--base_path", "b.fbin", "--query_path", "q.fbin", "--gt_path", "g.gt"])
    assert r.returncode == 1 and "float rows" in r.stderr
    r = run([tool("Vamana"), "--quant_base_path", "b.u8bin", "--index_path", "g",
             "--dir_bias_scale", "0.5"])
    assert r.returncode == 1 and "dir_bias_scale" in r.stderr


@pytest.mark.gpu
def test_compgt_tool_writes_the_reference_gt(gpu, golden, tmp_path):
    write_bin(tmp_path / "b.u8bin", golden["base"])
    write_bin(tmp_path / "q.u8bin", golden["queries"])
    for engine in ("mfma", "plain"):
        gt = tmp_path / ("g_%s.gt" % engine)
        r = run([tool("CompGT"), "--base_path", str(tmp_path / "b.u8bin"), "--query_path",
                 str(tmp_path / "q.u8bin"), "--gt_path", str(gt), "--k", "50", "--engine", engine])
        assert r.returncode == 0, r.stderr
        raw = np.fromfile(gt, dtype=np.uint32)
        nq, k = raw[0], raw[1]
        assert (nq, k) == (40, 50)
        assert (raw[2:2 + nq * k].reshape(nq, k) == golden["gt_ids"]).all()
        assert (raw[2 + nq * k:].view(np.float32).reshape(nq, k) == golden["gt_dists"]).all()


@pytest.mark.gpu
def test_vamana_tool_builds_the_reference_graph(gpu, golden, tmp_path):
    write_bin(tmp_path / "b.u8bin", golden["base"])
    write_bin(tmp_path / "q.u8bin", golden["queries"])
    r = run([tool("Vamana"), "--quant_base_path", str(tmp_path / "b.u8bin"), "--index_path",
             str(tmp_path / "g1"), "--R", "24", "--L", "48", "--max_alpha", "1.2", "--num_split",
             "0", "--num_passes", "1", "--rebuild"])
    assert r.returncode == 0, r.stderr
    assert (live(read_graph(tmp_path / "g1.graph")) == live(golden["graph1"])).all()
    # the recall sweep over the saved graph (no rebuild), reference-exact search
    gt = tmp_path / "g.gt"
    r = run([tool("CompGT"), "--base_path", str(tmp_path / "b.u8bin"), "--query_path",
             str(tmp_path / "q.u8bin"), "--gt_path", str(gt), "--k", "100"])
    assert r.returncode == 0, r.stderr
    csv = tmp_path / "res.csv"
    r = run([tool("Vamana"), "--quant_base_path", str(tmp_path / "b.u8bin"), "--index_path",
             str(tmp_path / "g1"), "--R", "24", "--quant_query_path", str(tmp_path / "q.u8bin"),
             "--gt_path", str(gt), "--res_path", str(csv), "--k", "10"])
    assert r.returncode == 0, r.stderr
    assert "For 10@10 recall" in r.stdout
    text = csv.read_text().splitlines()
    assert text[0].startswith('"GRAPH","Parameters"') and text[3].startswith('"Num queries"')
    rows = [ln.split(",") for ln in text[4:] if ln and ln[0].isdigit()]
    assert rows and all(int(x[0]) == 40 for x in rows)


def write_gt(path, ids, d):
    with open(path, "wb") as f:
        np.array(ids.shape, np.uint32).tofile(f)
        np.ascontiguousarray(ids, np.uint32).tofile(f)
        np.ascontiguousarray(d, np.float32).tofile(f)


def sweep_lines(out):
    """the per-point search lines of a --verbose sweep, without the timing"""
    return [ln.split(", QPS=")[0] for ln in out.splitlines() if ln.startswith("search: ")]


@pytest.mark.gpu
@pytest.mark.parametrize("name,ext,xext,rf", [("f32", "u8bin", "fbin", "100"),
                                              ("f32", "u8bin", "fbin", "0"),
                                              ("u8", "u4bin", "u8bin", "2")])
def test_vamana_tool_rerank_sweep_matches_reference_binary(gpu, golden_exact, tmp_path, name,
                                                           ext, xext, rf):
    """Quantized graph + raw rows: every point of the recall sweep (recall,
    mean visited, mean comparisons) equals the reference binary's."""
    import oracle as O
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "Vamana")
    if not os.path.exists(ref_bin):
        pytest.skip("oracle/_ref not built")
    z = golden_exact
    write_bin(tmp_path / ("b." + ext), z[name + "_codes"])
    write_bin(tmp_path / ("q." + ext), z[name + "_qcodes"])
    write_bin(tmp_path / ("x." + xext), z[name + "_raw"])
    write_bin(tmp_path / ("xq." + xext), z[name + "_rawq"])
    g = z[name + "_graph"]
    with open(tmp_path / "g.graph", "wb") as f:
        np.array([g.shape[0], g.shape[1] - 1], np.uint32).tofile(f)
        g[:, 0].tofile(f)
        for row in g:
            row[1:1 + row[0]].tofile(f)
    xdt = O.F32 if xext == "fbin" else O.U8
    gi, gd = O.compgt(z[name + "_raw"], z[name + "_rawq"], 100, xdt, tie_mode=1)
    write_gt(tmp_path / "g.gt", gi, gd)
    args = ["--quant_base_path", str(tmp_path / ("b." + ext)), "--index_path", str(tmp_path / "g"),
            "--quant_query_path", str(tmp_path / ("q." + ext)), "--base_path",
            str(tmp_path / ("x." + xext)), "--query_path", str(tmp_path / ("xq." + xext)),
            "--gt_path", str(tmp_path / "g.gt"), "--k", "10", "--rerank_factor", rf, "--verbose"]
    ours = run([tool("Vamana")] + args + ["--res_path", str(tmp_path / "ours.csv")])
    assert ours.returncode == 0, ours.stderr
    ref = subprocess.run([ref_bin] + args + ["--res_path", str(tmp_path / "ref.csv")],
                         capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr[-2000:]
    a, b = sweep_lines(ours.stdout), sweep_lines(ref.stdout)
    assert len(a) == len(b) > 60
    assert a == b


@pytest.mark.gpu
@pytest.mark.parametrize("name,ext", [("u8", "u8bin"), ("i8", "i8bin")])
def test_mips_tools_match_reference_binaries(gpu, golden_mips, tmp_path, name, ext):
    """--use_mips: CompGT writes the reference binary's .gt, Vamana builds its
    graph, and every point of the recall sweep equals the reference's."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "Vamana")
    if not os.path.exists(ref_bin):
        pytest.skip("oracle/_ref not built")
    z = golden_mips
    b, q = tmp_path / ("b." + ext), tmp_path / ("q." + ext)
    write_bin(b, z[name + "_base"])
    write_bin(q, z[name + "_queries"])
    gt = tmp_path / "g.gt"
    r = run([tool("CompGT"), "--base_path", str(b), "--query_path", str(q), "--gt_path", str(gt),
             "--k", "50", "--use_mips"])
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(gt, dtype=np.uint32)
    nq, k = int(raw[0]), int(raw[1])
    assert (raw[2:2 + nq * k].reshape(nq, k) == z[name + "_gt_ids"]).all()
    assert (raw[2 + nq * k:].view(np.float32).reshape(nq, k) == z[name + "_gt_dists"]).all()
    build = ["--quant_base_path", str(b), "--R", "24", "--L", "48", "--max_alpha", "1.2",
             "--num_split", "0", "--num_passes", "1", "--use_mips"]
    r = run([tool("Vamana"), "--index_path", str(tmp_path / "g"), "--rebuild"] + build)
    assert r.returncode == 0, r.stderr
    assert (live(read_graph(tmp_path / "g.graph")) == live(z[name + "_graph1"])).all()
    sweep = build + ["--index_path", str(tmp_path / "g"), "--quant_query_path", str(q),
                     "--gt_path", str(gt), "--k", "10", "--verbose"]
    ours = run([tool("Vamana")] + sweep + ["--res_path", str(tmp_path / "ours.csv")])
    assert ours.returncode == 0, ours.stderr
    ref = subprocess.run([ref_bin] + sweep + ["--res_path", str(tmp_path / "ref.csv")],
                         capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr[-2000:]
    a, c = sweep_lines(ours.stdout), sweep_lines(ref.stdout)
    assert len(a) == len(c) > 60 and a == c


@pytest.mark.gpu
def test_pcacompress_tool_writes_the_reference_codes(gpu, tmp_path):
    g = np.load(os.path.join(GOLD, "reference_pca.npz"))
    for t in "fg":
        write_bin(tmp_path / ("x_%s.fbin" % t), g["x_" + t])
        W = g["W_" + t]
        with open(tmp_path / ("w_%s.mat" % t), "wb") as f:
            np.array([W.shape[0], W.shape[1]], np.uint32).tofile(f)
            np.array([bool(g["hasbias_" + t][0])], np.uint8).tofile(f)
            W.tofile(f)
            if g["hasbias_" + t][0]:
                g["b_" + t].tofile(f)
        for ext in ("fbin", "u8bin", "i8bin", "u4bin", "i4bin"):
            out = tmp_path / ("y_%s.%s" % (t, ext))
            r = run([tool("PCACompress"), "--ifile", str(tmp_path / ("x_%s.fbin" % t)), "--ofile",
                     str(out), "--matFile", str(tmp_path / ("w_%s.mat" % t))])
            assert r.returncode == 0, r.stderr
            raw = np.fromfile(out, dtype=np.uint8)
            n, d = raw[:8].view(np.uint32)
            want = g["out_%s_%s" % (t, ext)]
            assert (n, d) == want.shape
            body = raw[8:]
            got = body.view(np.float32).reshape(n, d) if ext == "fbin" else body.reshape(n, d)
            assert (got == want).all(), (t, ext)
