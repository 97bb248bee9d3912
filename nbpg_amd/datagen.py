"""Seeded synthetic workloads for the NBPG hot path.

The reference's benchmarks use SIFT1M downloaded over ftp; offline, every
workload here is a seeded Gaussian-cluster mixture of the same shape.  All
draws come from numpy's PCG64 with fixed seeds, so the vectors are identical
on every host.
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def cluster_centers_u8(n_clusters, dims, seed, lo=20.0, hi=200.0):
    return _rng(seed).uniform(lo, hi, size=(n_clusters, dims)).astype(np.float32)


def mixture_u8(n, dims, n_clusters, sigma, seed, centers_seed=None, scale=None, shift=0.0,
               chunk=1 << 18):
    """u8 rows: round(scale * (center + sigma * N(0,1)) - shift) clipped to 0..255.

    `scale` is an optional per-dimension multiplier (heavy-tailed workloads)
    and `shift` moves mass below zero so clipping yields exact zeros; centres
    come from `centers_seed` so base and queries share them."""
    centers = cluster_centers_u8(n_clusters, dims, centers_seed if centers_seed is not None else 7)
    g = _rng(seed)
    out = np.empty((n, dims), np.uint8)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = g.integers(0, n_clusters, size=e - s)
        noise = g.standard_normal(size=(e - s, dims), dtype=np.float32) * np.float32(sigma)
        x = centers[lab] + noise
        if scale is not None:
            x = x * scale - np.float32(shift)
        np.clip(np.rint(x), 0, 255, out=x)
        out[s:e] = x.astype(np.uint8)
    return out


def heavy_tail_scale(dims, seed=11, sigma=0.5):
    """Per-dimension lognormal(0, sigma) scale used by the 1M x 128 workload."""
    return _rng(seed).lognormal(0.0, sigma, size=dims).astype(np.float32)


def mixture_i8(n, dims, n_clusters, sigma, seed, centers_seed=5, chunk=1 << 18):
    """i8 rows in -127..127 from a cluster mixture."""
    centers = _rng(centers_seed).uniform(-90.0, 90.0, size=(n_clusters, dims)).astype(np.float32)
    g = _rng(seed)
    out = np.empty((n, dims), np.int8)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = g.integers(0, n_clusters, size=e - s)
        x = centers[lab] + g.standard_normal(size=(e - s, dims), dtype=np.float32) * np.float32(sigma)
        np.clip(np.rint(x), -127, 127, out=x)
        out[s:e] = x.astype(np.int8)
    return out


def workload_small(n=100_000, nq=1_000, dims=128):
    """100k x 128 u8, 100 isotropic clusters (centres U[20,200], sigma 25)."""
    base = mixture_u8(n, dims, 100, 25.0, seed=1, centers_seed=7)
    queries = mixture_u8(nq, dims, 100, 25.0, seed=2, centers_seed=7)
    return base, queries


def workload_1m(n=1_000_000, nq=10_000, dims=128):
    """1M x 128 u8, 1000 clusters (sigma 25), lognormal(0, 0.5) per-dimension
    scale and a shift that leaves about 30% of the entries at zero."""
    sc = heavy_tail_scale(dims)
    base = mixture_u8(n, dims, 1000, 25.0, seed=3, centers_seed=9, scale=sc, shift=66.0)
    queries = mixture_u8(nq, dims, 1000, 25.0, seed=4, centers_seed=9, scale=sc, shift=66.0)
    return base, queries


def workload_1m_queries(nq=10_000, batch=0, dims=128):
    """Query batch `batch` of the 1M x 128 u8 workload: the same mixture as
    its base; batch 0 is workload_1m's own query set, later batches are
    fresh draws."""
    sc = heavy_tail_scale(dims)
    return mixture_u8(nq, dims, 1000, 25.0, seed=4 if batch == 0 else 1000 + batch, centers_seed=9,
                      scale=sc, shift=66.0)


def workload_serve(n=10_000_000, nq=10_000, dims=96, seed=21):
    """10M x 96 i8 rows in -127..127: 4096 clusters (centres U[-90, 90]),
    sigma 30.  Drawn on the device (torch's seeded generator) and returned
    as host arrays; base and queries share the centres."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    centers = torch.rand((4096, dims), generator=g, device="cuda") * 180.0 - 90.0

    def draw(m):
        out = torch.empty((m, dims), dtype=torch.int8, device="cuda")
        for s in range(0, m, 1 << 21):
            e = min(m, s + (1 << 21))
            lab = torch.randint(0, 4096, (e - s,), generator=g, device="cuda")
            x = centers[lab] + torch.randn((e - s, dims), generator=g, device="cuda") * 30.0
            out[s:e] = torch.clamp(torch.round(x), -127, 127).to(torch.int8)
        return out

    base = draw(n)
    queries = draw(nq)
    return base.cpu().numpy(), queries.cpu().numpy()


def workload_pca(n=1_000_000, nq=10_000, dims=960, clusters=1000, seed=31):
    """f32 rows from a Gaussian mixture with a decaying spectrum: in a random
    orthonormal basis Q, per-coordinate variance i^-1.2 (i = 1..dims), with
    cluster centres drawn from the same spectrum (scale 2).  Returned on the
    device: (base n x dims, queries nq x dims)."""
    import torch
    rng = _rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((dims, dims)))
    sd = (np.arange(1, dims + 1, dtype=np.float64) ** -0.6)
    B = torch.from_numpy((Q * sd[None, :]).T.astype(np.float32)).cuda()  # z @ B = (z*sd) Q^T
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    centers = 2.0 * (torch.randn((clusters, dims), generator=g, device="cuda") @ B)

    def draw(m):
        out = torch.empty((m, dims), dtype=torch.float32, device="cuda")
        for s in range(0, m, 1 << 18):
            e = min(m, s + (1 << 18))
            lab = torch.randint(0, clusters, (e - s,), generator=g, device="cuda")
            out[s:e] = centers[lab] + torch.randn((e - s, dims), generator=g, device="cuda") @ B
        return out

    return draw(n), draw(nq)
