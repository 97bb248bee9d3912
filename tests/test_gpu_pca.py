"""PCACompress on the device against the reference's own outputs, the oracle
and an FP64 projection."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_pca.npz")
CODES = [(1, 0, 1, "u8bin"), (2, 1, 1, "i8bin"), (3, 0, 2, "u4bin"), (4, 1, 2, "i4bin")]


def test_plain_projection_matches_reference_golden(gpu):
    """NBPG_PCA_PLAIN: y bit-identical to the reference's PCACompress .fbin
    output, codes identical for every code type."""
    N = gpu
    import torch
    g = np.load(GOLD)
    for t in "fg":
        x = torch.from_numpy(g["x_" + t]).cuda()
        W = torch.from_numpy(g["W_" + t]).cuda()
        b = torch.from_numpy(g["b_" + t]).cuda()
        _, y = N.pca_project(x, W, b, code=N.CODE_F32, mode=N.PCA_PLAIN)
        assert (y.cpu().numpy() == g["out_%s_fbin" % t]).all(), t
        for code, _, _, ext in CODES:
            c, _ = N.pca_project(x, W, b, code=code, mode=N.PCA_PLAIN)
            assert (c.cpu().numpy() == g["out_%s_%s" % (t, ext)]).all(), (t, ext)


def test_quantizer_bit_exact(gpu):
    """compress (matrix.h:99-127) on boundary values: -1, 1, just outside,
    exact (v+1)*127.5 integers, huge, infinite and NaN inputs."""
    N = gpu
    import torch
    rng = np.random.default_rng(3)
    y = (rng.standard_normal((257, 64)) * 0.8).astype(np.float32)
    special = np.array([-1.0, 1.0, -1.0000001, 1.0000001, 0.0, -0.0, 3e9, -3e9, np.inf, -np.inf,
                        np.nan, 2 / 255 - 1, 4 / 255 - 1, 1 - 1 / 255, 127 / 127.5 - 1,
                        128 / 127.5 - 1], np.float32)
    y[0, :len(special)] = special
    y[1, :] = np.linspace(-1.01, 1.01, 64, dtype=np.float32)
    yd = torch.from_numpy(y).cuda()
    for code, sg, ch, _ in CODES:
        got = N.pca_quantize(yd, code).cpu().numpy()
        assert (got == O.compress(y, sg, ch)).all(), code


@pytest.mark.parametrize("xt", ["f32", "u8", "i8"])
def test_plain_projection_matches_oracle(gpu, xt):
    N = gpu
    import torch
    rng = np.random.default_rng(11)
    n, d_in, d_out = 1500, 200, 24
    if xt == "f32":
        x = (rng.standard_normal((n, d_in)) * 0.1).astype(np.float32)
        xf = x
    else:
        x = rng.integers(0, 256, (n, d_in), dtype=np.uint8)
        if xt == "i8":
            x = x.view(np.int8)
        xf = x.astype(np.float32)
    W = (rng.standard_normal((d_out, d_in)) / (60.0 if xt == "f32" else 6000.0)).astype(np.float32)
    b = (rng.standard_normal(d_out) * 0.1).astype(np.float32)
    c, y = N.pca_project(torch.from_numpy(x).cuda(), torch.from_numpy(W).cuda(),
                         torch.from_numpy(b).cuda(), code=N.CODE_I4, mode=N.PCA_PLAIN,
                         want_y=True)
    yo = O.matmul(xf, W, b)
    assert (y.cpu().numpy() == yo).all()
    assert (c.cpu().numpy() == O.compress(yo, 1, 2)).all()


@pytest.mark.parametrize("xt,d_in,d_out", [("f32", 960, 64), ("f32", 136, 16), ("u8", 128, 32),
                                           ("i8", 96, 128)])
def test_bf16_projection_error_bound(gpu, xt, d_in, d_out):
    """NBPG_PCA_BF16 against an FP64 projection.  Each operand is rounded to
    bf16 (relative error <= 2^-9), products are exact in fp32 and sums carry
    at most d_in * 2^-24 relative error, so |y - y64| <= 2^-7 * sum|w||x|
    holds with a 2x margin.  The fused codes equal the quantizer applied to
    the kernel's own y."""
    N = gpu
    import torch
    rng = np.random.default_rng(d_in + d_out)
    n = 4099
    if xt == "f32":
        x = (rng.standard_normal((n, d_in)) * 0.05).astype(np.float32)
        xf = x.astype(np.float64)
    else:
        x = rng.integers(0, 256, (n, d_in), dtype=np.uint8)
        if xt == "i8":
            x = x.view(np.int8)
        xf = x.astype(np.float64)
    W = (rng.standard_normal((d_out, d_in)) / (8.0 * np.sqrt(d_in) * (1 if xt == "f32" else 120))
         ).astype(np.float32)
    b = (rng.standard_normal(d_out) * 0.05).astype(np.float32)
    xd, Wd, bd = (torch.from_numpy(a).cuda() for a in (x, W, b))
    y64 = xf @ W.astype(np.float64).T + b
    bound = np.abs(xf) @ np.abs(W.astype(np.float64)).T * 2.0 ** -7 + 1e-6
    for code, sg, ch, _ in CODES:
        c, y = N.pca_project(xd, Wd, bd, code=code, mode=N.PCA_BF16, want_y=True)
        y = y.cpu().numpy()
        err = np.abs(y - y64)
        assert (err <= bound).all(), (code, err.max())
        assert (c.cpu().numpy() == O.compress(y, sg, ch)).all(), code
    print("bf16 %s %dx%d: max abs err %.3g, mean %.3g" % (xt, d_in, d_out, err.max(), err.mean()))


@pytest.mark.parametrize("xt,d_in,d_out,code", [("f32", 96, 64, "CODE_I8"), ("u8", 64, 32, "CODE_U4"),
                                                ("f32", 960, 64, "CODE_I4")])
def test_bf16_projection_lds_path(gpu, xt, d_in, d_out, code):
    """Large inputs take the persistent kernel with W in LDS; every 64-row
    block is computed as by the global-W kernel (same operands, same MFMA
    order), so its rows equal those of a small call over the same rows, bit
    for bit, for y and the fused codes."""
    N = gpu
    import torch
    rng = np.random.default_rng(d_in)
    n = 420_000 if d_in < 960 else 410_000
    if xt == "f32":
        x = torch.randn((n, d_in), device="cuda", dtype=torch.float32) * 0.05
    else:
        x = torch.randint(0, 256, (n, d_in), device="cuda", dtype=torch.uint8)
    W = torch.from_numpy((rng.standard_normal((d_out, d_in)) / (8.0 * np.sqrt(d_in) * (1 if xt == "f32" else 120))
                          ).astype(np.float32)).cuda()
    b = torch.from_numpy((rng.standard_normal(d_out) * 0.05).astype(np.float32)).cuda()
    cd = getattr(N, code)
    c, y = N.pca_project(x, W, b, code=cd, mode=N.PCA_BF16, want_y=True)
    for lo in (0, 123_456, n - 1000):
        cs, ys = N.pca_project(x[lo:lo + 1000].contiguous(), W, b, code=cd, mode=N.PCA_BF16,
                               want_y=True)
        # a row's MFMA sums do not depend on the other rows of its block
        assert torch.equal(ys, y[lo:lo + 1000]) and torch.equal(cs, c[lo:lo + 1000]), lo


def test_stats_f64(gpu):
    """FP64 mean and covariance against numpy's float64, d not a multiple of
    the 64-column tiles, f32 and u8 rows."""
    N = gpu
    import torch
    rng = np.random.default_rng(5)
    for x in [(rng.standard_normal((7001, 100)) * np.linspace(0.1, 3, 100)).astype(np.float32),
              rng.integers(0, 256, (3000, 70), dtype=np.uint8)]:
        mean, cov = N.pca_stats(torch.from_numpy(x).cuda())
        xf = x.astype(np.float64)
        m64 = xf.mean(axis=0)
        c64 = np.cov(xf, rowvar=False)
        assert np.allclose(mean.cpu().numpy(), m64, rtol=1e-12, atol=1e-12)
        c = cov.cpu().numpy()
        assert (c == c.T).all()
        assert np.abs(c - c64).max() <= 1e-10 * np.abs(c64).max()


def test_fit_recovers_subspace(gpu):
    """Rows on a 3-dimensional subspace plus small noise: the fit's first
    three directions span it, the projection of a sample lies in [-1, 1]."""
    N = gpu
    import torch
    from nbpg_amd import pca as P
    rng = np.random.default_rng(8)
    Q, _ = np.linalg.qr(rng.standard_normal((40, 3)))
    z = rng.standard_normal((20000, 3)) * np.array([5.0, 3.0, 2.0])
    x = (z @ Q.T + 0.01 * rng.standard_normal((20000, 40)) + 7.0).astype(np.float32)
    f = P.fit(torch.from_numpy(x).cuda(), 3)
    V = f["W"].astype(np.float64) * f["scale"]
    assert np.allclose(V @ V.T, np.eye(3), atol=1e-5)
    assert np.linalg.norm(V @ Q @ Q.T - V) < 1e-3
    assert f["eigenvalues"][0] > f["eigenvalues"][1] > f["eigenvalues"][2] > 100 * f["eigenvalues"][3]
    _, y = N.pca_project(torch.from_numpy(x).cuda(), torch.from_numpy(f["W"]).cuda(),
                         torch.from_numpy(f["bias"]).cuda(), code=N.CODE_F32, mode=N.PCA_PLAIN)
    assert np.abs(y.cpu().numpy()).max() <= 1.0 + 1e-5


def test_fit_clip_quantile(gpu):
    """fit(clip=q) maps the q-quantile of the sample's |projection| to 1: a
    fraction 1 - q of the sample's projected values lands outside [-1, 1]
    (the quantizer saturates them), the directions are the clip = 1 fit's,
    and the scale only shrinks."""
    N = gpu
    import torch
    from nbpg_amd import pca as P
    rng = np.random.default_rng(9)
    x = (rng.standard_normal((30000, 48)) * np.linspace(3, 0.2, 48)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    f1 = P.fit(xd, 16, sample=30000)
    for q in (0.999, 0.99, 0.95):
        f = P.fit(xd, 16, sample=30000, clip=q)
        assert f["scale"] < f1["scale"]
        assert np.allclose(f["W"] * f["scale"], f1["W"] * f1["scale"], atol=1e-6)
        y = (x.astype(np.float64) @ f["W"].astype(np.float64).T + f["bias"])
        out = float((np.abs(y) > 1.0).mean())
        assert abs(out - (1 - q)) < 0.002 + 0.1 * (1 - q), (q, out)
