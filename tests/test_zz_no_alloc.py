"""nbpg_search_rerank_noalloc (scratch in the library's fixed per-GPU array,
no device allocation) returns exactly what nbpg_search_rerank returns: ids,
distances and {visited, dist_cmps} counters, bit for bit, in FAST mode (nq x
L candidate ids, or nq visited lists, in the scratch) and in reference mode at an L whose search
state does not fit in LDS (per-workgroup state in the scratch, a smaller
grid).  Each case runs two calls in a row, the second with fewer queries,
so the second call reads a scratch that still holds the first call's
contents."""
import numpy as np
import pytest

from nbpg_amd import datagen as D

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20  # NBPG_NOALLOC_SCRATCH_BYTES in include/nbpg.h


@pytest.fixture(scope="module")
def setup(gpu):
    N = gpu
    import torch
    rng = np.random.default_rng(21)
    raw = (rng.standard_normal((8000, 64)) * np.linspace(2, 0.3, 64)).astype(np.float32)
    rq = (rng.standard_normal((400, 64)) * np.linspace(2, 0.3, 64)).astype(np.float32)
    codes = np.clip(np.rint(raw * 20 + 128), 0, 255).astype(np.uint8)
    qcodes = np.clip(np.rint(rq * 20 + 128), 0, 255).astype(np.uint8)
    ix = N.Index(N.U8, len(codes), codes.shape[1], 32)
    ix.set_points(codes)
    N.build(ix, 32, 64, alpha=1.2, passes=1, num_split=0)
    qpad = np.zeros((len(qcodes), ix.row_stride), np.uint8)
    qpad[:, :qcodes.shape[1]] = qcodes
    yield N, ix, N.to_device(qpad), torch.from_numpy(raw).cuda(), torch.from_numpy(rq).cuda()
    ix.close()


def _same(a, b):
    import torch
    return all(torch.equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("mode,L,k,rf,sizes", [
    ("fast", 64, 10, 4, (400, 150)),
    ("fast", 200, 10, 100, (300, 37)),
    ("fast", 32, 10, 0, (400, 90)),    # filtering: visited lists in the scratch
    ("ref", 256, 10, 2, (200, 60)),    # state beyond LDS: global scratch, smaller grid
    ("ref", 256, 10, 0, (120, 31)),    # filtering mode
    ("ref", 48, 10, 3, (200, 90)),     # state in LDS
])
def test_noalloc_matches_allocating_entry(setup, mode, L, k, rf, sizes):
    N, ix, qd, xb, xq = setup
    m = N.SEARCH_FAST if mode == "fast" else N.SEARCH_REF
    for nq in sizes:
        want = N.search_rerank(ix, qd[:nq], xb, xq[:nq], L, k, rf, mode=m)
        got = N.search_rerank(ix, qd[:nq], xb, xq[:nq], L, k, rf, mode=m, noalloc=True)
        assert _same(got, want), (mode, L, k, rf, nq)
        assert (want[0].cpu().numpy().view(np.uint32) < ix.n).all()


def test_noalloc_rejects_sizes_beyond_scratch(setup):
    N, ix, qd, xb, xq = setup
    import torch
    L = 256
    nq = SCRATCH_BYTES // (4 * L) + 1
    big_q = qd[:1].expand(nq, -1).contiguous()
    big_rq = xq[:1].expand(nq, -1).contiguous()
    with pytest.raises(N.NbpgError, match="scratch"):
        N.search_rerank(ix, big_q, xb, big_rq, L, 10, 4, mode=N.SEARCH_FAST, noalloc=True)
    # the largest FAST call that fits still runs, and equals the allocating entry
    nq = SCRATCH_BYTES // (4 * L)
    small_q, small_rq = big_q[:nq], big_rq[:nq]
    got = N.search_rerank(ix, small_q, xb, small_rq, L, 10, 4, mode=N.SEARCH_FAST, noalloc=True)
    want = N.search_rerank(ix, small_q, xb, small_rq, L, 10, 4, mode=N.SEARCH_FAST)
    torch.cuda.synchronize()
    assert _same(got, want)
