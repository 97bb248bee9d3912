"""Write tests/golden/reference_pca.npz from the reference's own PCACompress.

Runs only in the container that holds the reference tree: oracle/Makefile
builds the reference's PCACompress.cpp into oracle/_ref/PCACompress, which is
run on small .fbin / .u8bin / .i8bin inputs and .mat files written here, for
every output extension it dispatches on (PCACompress.cpp:25-37).  The
fixture keeps the inputs and the files' payloads; no reference text.

    python tests/golden/make_golden_pca.py
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

EXTS = [".fbin", ".u8bin", ".i8bin", ".u4bin", ".i4bin"]


def write_bin(path, a):
    a = np.ascontiguousarray(a)
    with open(path, "wb") as f:
        np.array([a.shape[0], a.shape[1]], np.uint32).tofile(f)
        a.tofile(f)


def write_mat(path, W, bias):
    """[u32 d_out][u32 d_in][u8 has_bias][f32 W][f32 bias] (matrix.h:59-83)."""
    with open(path, "wb") as f:
        np.array([W.shape[0], W.shape[1]], np.uint32).tofile(f)
        np.array([bias is not None], np.uint8).tofile(f)
        W.astype(np.float32).tofile(f)
        if bias is not None:
            bias.astype(np.float32).tofile(f)


def read_out(path, ext):
    raw = np.fromfile(path, dtype=np.uint8)
    n, d = raw[:8].view(np.uint32)
    body = raw[8:]
    if ext == ".fbin":
        return n, d, body.view(np.float32).reshape(n, d)
    return n, d, body.reshape(n, d)


def main():
    O.build(ref=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "PCACompress")
    assert os.path.exists(exe), "oracle/_ref/PCACompress not built"
    tmp = tempfile.mkdtemp()
    rng = np.random.default_rng(77)
    out = {}
    # f32 rows with a bias-carrying .mat; y spread over about [-1.3, 1.3]
    x = (rng.standard_normal((300, 96)) * 0.12).astype(np.float32)
    W = rng.standard_normal((16, 96)).astype(np.float32)
    W /= np.linalg.norm(W, axis=1, keepdims=True)
    b = (rng.standard_normal(16) * 0.2).astype(np.float32)
    # a .mat without bias.  (The reference's .u8bin / .i8bin inputs are not
    # recorded: its loader keeps a 100000-row block that is never resized,
    # PCACompress.cpp:52-61 with matrix.h:36-55, and aborts unless n is a
    # multiple of 100000.)
    x2 = (rng.standard_normal((128, 64)) * 0.2).astype(np.float32)
    W2 = (rng.standard_normal((32, 64)) * 0.15).astype(np.float32)
    cases = [("f", x, ".fbin", W, b), ("g", x2, ".fbin", W2, None)]
    for tag, xs, iext, Wm, bias in cases:
        ipath = os.path.join(tmp, "x_" + tag + iext)
        mpath = os.path.join(tmp, "w_" + tag + ".mat")
        write_bin(ipath, xs)
        write_mat(mpath, Wm, bias)
        out["x_" + tag] = xs
        out["W_" + tag] = Wm
        out["b_" + tag] = bias if bias is not None else np.zeros(Wm.shape[0], np.float32)
        out["hasbias_" + tag] = np.array([bias is not None])
        for ext in EXTS:
            opath = os.path.join(tmp, "y_" + tag + ext)
            subprocess.run([exe, "--ifile", ipath, "--ofile", opath, "--matFile", mpath],
                           check=True, stdout=subprocess.DEVNULL)
            n, d, payload = read_out(opath, ext)
            assert n == xs.shape[0]
            out["out_%s_%s" % (tag, ext[1:])] = payload
    np.savez_compressed(os.path.join(HERE, "reference_pca.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_pca.npz"))


if __name__ == "__main__":
    main()
