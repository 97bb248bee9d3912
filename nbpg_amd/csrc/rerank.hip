// rerank.hip -- the rerank step of beam_search_rerank (beamSearch.h:183-236):
// a search over quantized codes returns its frontier; the first
// num_check = min(k * rerank_factor, frontier size) entries are re-scored with
// the full-precision rows, sorted by (dist, id) and cut to k.
//
// One wavefront per query.  Each lane scores whole candidates: it streams its
// candidate's raw row in index order, so a float distance is the reference's
// sequential fp32 sum (point.h:137-146) bit for bit and an integer one is
// exact.  The <= 1024 keys are then sorted in LDS.
#include <algorithm>

#include "raw_rows.h"

namespace nbpg {

struct RerankArgs {
    const uint8_t *base, *queries;
    uint64_t n, nq, row_bytes, qstride;
    uint32_t dims, cand_stride, num_check, k;
    const uint32_t *cand;
    uint32_t *ids;
    float *dists;
};

template <int RT>
__global__ __launch_bounds__(256) void rerank_kernel(RerankArgs a)
{
    __shared__ uint64_t keys[4][RERANK_MAX];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint64_t q = (uint64_t)blockIdx.x * 4 + wv;
    if (q >= a.nq)
        return;
    uint64_t *kw = keys[wv];
    const uint8_t *qrow = a.queries + q * a.qstride;
    // the frontier ends at its first empty slot
    uint32_t m = a.num_check;
    for (uint32_t i = lane; i < a.num_check; i += WAVE) {
        const uint32_t id = a.cand[q * a.cand_stride + i];
        if (id >= a.n)
            m = min(m, i);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = min(m, (uint32_t)__shfl_xor(m, o));
    uint32_t n2 = 1;
    while (n2 < m)
        n2 <<= 1;
    for (uint32_t i = lane; i < n2; i += WAVE) {
        uint64_t key = ~0ull;
        if (i < m) {
            const uint32_t id = a.cand[q * a.cand_stride + i];
            key = make_key(raw_row<RT>::dist(a.base + (uint64_t)id * a.row_bytes, qrow, a.dims), id);
        }
        kw[i] = key;
    }
    wave_sync();
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n2; i += WAVE) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = kw[i], y = kw[p];
                    if ((x > y) == ((i & kk) == 0)) {
                        kw[i] = y;
                        kw[p] = x;
                    }
                }
            }
            wave_sync();
        }
    for (uint32_t i = lane; i < a.k; i += WAVE) {
        const bool ok = i < m;
        a.ids[q * a.k + i] = ok ? key_id(kw[i]) : NBPG_SENT;
        if (a.dists)
            a.dists[q * a.k + i] = ok ? key_dist(kw[i]) : __int_as_float(0x7f800000);
    }
}

int rerank_launch(int raw_dtype, const void *d_raw_base, uint64_t n, uint32_t dims,
                  const void *d_raw_queries, uint64_t q_stride, uint64_t nq,
                  const uint32_t *d_cand, uint32_t cand_stride, uint32_t n_cand, uint32_t k,
                  uint32_t rerank_factor, uint32_t *d_out_ids, float *d_out_dists,
                  hipStream_t s)
{
    const uint64_t nc = std::min<uint64_t>((uint64_t)k * rerank_factor, n_cand);
    if (nc > RERANK_MAX)
        return set_error(NBPG_EUNSUP, "%llu candidates to re-score (at most %d)",
                         (unsigned long long)nc, RERANK_MAX);
    if (nq == 0)
        return NBPG_OK;
    RerankArgs a;
    a.base = (const uint8_t *)d_raw_base;
    a.queries = (const uint8_t *)d_raw_queries;
    a.n = n;
    a.nq = nq;
    a.row_bytes = (uint64_t)dims * raw_elem_bytes(raw_dtype);
    a.qstride = q_stride;
    a.dims = dims;
    a.cand_stride = cand_stride;
    a.num_check = (uint32_t)nc;
    a.k = k;
    a.cand = d_cand;
    a.ids = d_out_ids;
    a.dists = d_out_dists;
    const dim3 grid((uint32_t)((nq + 3) / 4)), block(256);
    if (raw_dtype == NBPG_F32)
        hipLaunchKernelGGL(rerank_kernel<NBPG_F32>, grid, block, 0, s, a);
    else if (raw_dtype == NBPG_U8)
        hipLaunchKernelGGL(rerank_kernel<NBPG_U8>, grid, block, 0, s, a);
    else if (raw_dtype == NBPG_I8)
        hipLaunchKernelGGL(rerank_kernel<NBPG_I8>, grid, block, 0, s, a);
    else
        hipLaunchKernelGGL(rerank_kernel<NBPG_U4X2>, grid, block, 0, s, a);
    return hipGetLastError() == hipSuccess ? NBPG_OK : set_error(NBPG_EHIP, "rerank launch");
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_rerank(int raw_dtype, const void *d_raw_base, uint64_t n, uint32_t dims,
                           const void *d_raw_queries, uint64_t nq, const uint32_t *d_cand,
                           uint32_t cand_stride, uint32_t n_cand, uint32_t k,
                           uint32_t rerank_factor, uint32_t *d_out_ids, float *d_out_dists,
                           void *stream)
{
    NBPG_TRY(ensure_init());
    if (raw_dtype < NBPG_U8 || raw_dtype > NBPG_F32)
        return set_error(NBPG_EINVAL, "raw dtype %d", raw_dtype);
    if (nq && (!d_raw_base || !d_raw_queries || !d_cand || !d_out_ids))
        return set_error(NBPG_EINVAL, "null argument");
    if (k == 0 || rerank_factor == 0 || n_cand == 0 || cand_stride < n_cand || dims == 0)
        return set_error(NBPG_EINVAL, "k, rerank_factor and n_cand must be positive");
    return rerank_launch(raw_dtype, d_raw_base, n, dims, d_raw_queries,
                         (uint64_t)dims * raw_elem_bytes(raw_dtype), nq, d_cand, cand_stride,
                         n_cand, k, rerank_factor, d_out_ids, d_out_dists, as_stream(stream));
}
