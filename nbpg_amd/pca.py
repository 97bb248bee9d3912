"""PCA fit, .mat files and file-level PCACompress for the device projection.

The reference names its fit script (py_data_tools/pca_process.py,
README.md:43) but does not ship it.  The fit here: FP64 mean and covariance
on the device (nbpg_pca_stats_f64), FP64 symmetric eigendecomposition on the
host (numpy.linalg.eigh, LAPACK), the d_out leading eigenvectors as the rows
of W, bias = -W mean, and one global scale so that the projections of a
sample fall in [-1, 1], the quantizer's input range (matrix.h:99-127).  A
single scale keeps the code space a scaled copy of the PCA space, so L2
neighbours are preserved up to quantization.
"""
import numpy as np

import nbpg_amd as N


def save_mat(path, W, bias=None):
    """.mat: [u32 d_out][u32 d_in][u8 has_bias][f32 W d_out x d_in][f32 bias]
    (loadMatrix, matrix.h:59-83)."""
    W = np.ascontiguousarray(W, dtype=np.float32)
    with open(path, "wb") as f:
        np.array([W.shape[0], W.shape[1]], np.uint32).tofile(f)
        np.array([bias is not None], np.uint8).tofile(f)
        W.tofile(f)
        if bias is not None:
            np.ascontiguousarray(bias, dtype=np.float32).tofile(f)


def load_mat(path):
    """(W, bias); bias is zeros when the file has none (matrix.h:76-80)."""
    raw = np.fromfile(path, dtype=np.uint8)
    d_out, d_in = (int(v) for v in raw[:8].view(np.uint32))
    has_bias = bool(raw[8])
    body = raw[9:]
    W = body[:4 * d_out * d_in].view(np.float32).reshape(d_out, d_in).copy()
    if has_bias:
        bias = body[4 * d_out * d_in:4 * d_out * (d_in + 1)].view(np.float32).copy()
    else:
        bias = np.zeros(d_out, np.float32)
    return W, bias


def fit(x_dev, d_out, sample=100000, margin=1.0, clip=1.0):
    """Returns dict(W f32 d_out x d_in, bias f32, eigenvalues f64 (all,
    descending), scale, mean f64).

    The scale maps the `clip` quantile of |projection| over a sample (all
    components together) to 1: clip = 1 takes the largest value, so every
    projection of the sample lies in [-1, 1]; a smaller clip lets the
    quantizer saturate the largest values (compress clamps them,
    matrix.h:99-127) in exchange for finer steps everywhere else, which
    4-bit codes need.  One global scale either way, so the code space stays
    a scaled copy of the PCA space."""
    import torch
    mean, cov = N.pca_stats(x_dev)
    torch.cuda.synchronize()
    cov_h = cov.cpu().numpy()
    evals, evecs = np.linalg.eigh(cov_h)
    order = np.argsort(evals)[::-1]
    evals = evals[order]
    V = evecs[:, order[:d_out]].T.copy()  # d_out x d_in, FP64
    # fixed sign per component: its largest-magnitude loading positive
    piv = np.abs(V).argmax(axis=1)
    V *= np.sign(V[np.arange(d_out), piv])[:, None]
    mean_h = mean.cpu().numpy()
    xs = x_dev[:sample].to(torch.float64)
    proj = (xs - torch.from_numpy(mean_h).to(xs.device)) @ torch.from_numpy(V.T.copy()).to(xs.device)
    a = proj.abs().flatten()
    if clip >= 1.0:
        scale = float(a.max().item()) * margin
    else:
        k = max(1, int(round(clip * a.numel())))
        scale = float(a.kthvalue(k).values.item()) * margin
    scale = scale if scale > 0 else 1.0
    W = V / scale
    bias = -(W @ mean_h)
    return {"W": W.astype(np.float32), "bias": bias.astype(np.float32), "eigenvalues": evals,
            "scale": scale, "mean": mean_h}


def read_bin(path):
    """.fbin / .u8bin / .i8bin: [u32 n][u32 dims][rows] (point_range.h:17-30)."""
    raw = np.fromfile(path, dtype=np.uint8)
    n, d = (int(v) for v in raw[:8].view(np.uint32))
    body = raw[8:]
    if path.endswith((".fbin", ".bin")):
        return body[:4 * n * d].view(np.float32).reshape(n, d)
    if path.endswith(".i8bin"):
        return body[:n * d].view(np.int8).reshape(n, d)
    return body[:n * d].reshape(n, d)


def compress_file(ifile, ofile, mat_file, mode=N.PCA_BF16, block=1000000):
    """PCACompress --ifile --ofile --matFile (PCACompress.cpp:21-79): the
    output extension picks the code type; rows stream through the device in
    blocks."""
    import torch
    W, bias = load_mat(mat_file)
    ext = "." + ofile.rsplit(".", 1)[-1]
    code = N.CODE_OF_EXT[ext]
    chunk = 2 if code in (N.CODE_U4, N.CODE_I4) else 1
    x = read_bin(ifile)
    n = x.shape[0]
    W_d = torch.from_numpy(W).cuda()
    b_d = torch.from_numpy(bias).cuda()
    with open(ofile, "wb") as f:
        np.array([n, W.shape[0] // chunk], np.uint32).tofile(f)
        for s in range(0, n, block):
            xb = torch.from_numpy(np.ascontiguousarray(x[s:s + block])).cuda()
            codes, y = N.pca_project(xb, W_d, b_d, code=code, mode=mode)
            out = y if code == N.CODE_F32 else codes
            out.cpu().numpy().tofile(f)
