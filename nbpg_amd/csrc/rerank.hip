// rerank.hip -- the rerank step of beam_search_rerank (beamSearch.h:183-236):
// a search over quantized codes returns its frontier; the first
// num_check = min(k * rerank_factor, frontier size) entries are re-scored with
// the full-precision rows, sorted by (dist, id) and cut to k.
//
// One wavefront per query.  Each lane scores whole candidates, summing its
// candidate's raw row in index order, so a float distance is the reference's
// sequential fp32 sum (point.h:137-146) bit for bit and an integer one is
// exact.  Aligned rows are staged through LDS with whole-line loads (the
// staged kernel); others are read lane by lane.  The <= 1024 keys are then
// sorted in LDS.
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

// frontier length of query q: its first id >= n ends it (<= num_check)
__device__ __forceinline__ uint32_t frontier_len(const RerankArgs &a, uint64_t q, uint32_t lane)
{
    uint32_t m = a.num_check;
    for (uint32_t i = lane; i < a.num_check; i += WAVE) {
        const uint32_t id = a.cand[q * a.cand_stride + i];
        if (id >= a.n)
            m = min(m, i);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = min(m, (uint32_t)__shfl_xor(m, o));
    return m;
}

// kw[0:n2) holds m keys then ~0: bitonic sort, then the first k out
__device__ __forceinline__ void sort_and_write(const RerankArgs &a, uint64_t q, uint64_t *kw,
                                               uint32_t m, uint32_t n2, uint32_t lane)
{
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

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t m)
{
    uint32_t n2 = 1;
    while (n2 < m)
        n2 <<= 1;
    return n2;
}

// any row size or alignment: each lane streams whole rows from memory
template <int RT, bool IP>
__global__ __launch_bounds__(256) void rerank_kernel(RerankArgs a)
{
    __shared__ uint64_t keys[4][RERANK_MAX];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint64_t q = (uint64_t)blockIdx.x * 4 + wv;
    if (q >= a.nq)
        return;
    uint64_t *kw = keys[wv];
    const uint8_t *qrow = a.queries + q * a.qstride;
    const uint32_t m = frontier_len(a, q, lane), n2 = pow2_at_least(m);
    for (uint32_t i = lane; i < n2; i += WAVE) {
        uint64_t key = ~0ull;
        if (i < m) {
            const uint32_t id = a.cand[q * a.cand_stride + i];
            key = make_key(raw_row<RT, IP>::dist(a.base + (uint64_t)id * a.row_bytes, qrow, a.dims), id);
        }
        kw[i] = key;
    }
    sort_and_write(a, q, kw, m, n2, lane);
}

// 16-byte aligned rows (row bytes % 16 == 0): up to 64 candidates at a time,
// their rows staged through LDS a piece per row per step with whole-line
// loads (one load instruction covers 8 rows x 128 B, or fewer rows with
// longer pieces), then each lane sums its own row's piece from LDS in index
// order -- the same sequential sum as above.  Row pitch in LDS: the piece
// plus a 16-byte pad against bank conflicts.
#define RS_STAGE (65 * 144) // 64 candidate rows of 128-byte pieces + the query's piece

typedef uint32_t rs_vec __attribute__((ext_vector_type(4))); // a 16-byte piece

// one step's loads: lane (sub, piece) takes piece `piece` (16 B) of rows
// t * (64 / PPR) + sub, t = 0..7, and lanes 0..PPR-1 the query's pieces
template <int PPR>
__device__ __forceinline__ void rs_fetch(rs_vec (&buf)[8], rs_vec &qbuf, const uint8_t *base,
                                         uint64_t row_bytes, uint32_t my_id, const uint8_t *qrow,
                                         uint64_t c0, uint32_t gm, uint32_t piece, uint32_t sub)
{
    constexpr uint32_t RPI = WAVE / PPR; // rows per load instruction
    const uint64_t off = c0 + 16u * piece;
    const bool in_row = off < row_bytes;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const uint32_t id = (uint32_t)__shfl((int)my_id, t * RPI + (int)sub);
        if ((uint32_t)t * RPI + sub < gm && in_row)
            buf[t] = *reinterpret_cast<const rs_vec *>(base + (uint64_t)id * row_bytes + off);
    }
    if (sub == 0 && in_row)
        qbuf = *reinterpret_cast<const rs_vec *>(qrow + off);
}

// PPR 16-byte pieces of each row per step: G = 8 * 64 / PPR candidates per
// pass (64 with PPR 8; few candidates take longer pieces, so a step still
// moves 8 KB and a row takes fewer dependent steps)
template <int RT, bool IP, int PPR>
__global__ __launch_bounds__(256) void rerank_staged_kernel(RerankArgs a, uint32_t n2cap)
{
    constexpr uint32_t RPI = WAVE / PPR, G = 8 * RPI, CH = 16 * PPR, PITCH = CH + 16;
    static_assert((G + 1) * PITCH <= RS_STAGE, "stage rows fit the per-wave stage");
    extern __shared__ __align__(16) uint8_t lds_rr[];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint64_t q = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + wv;
    if (q >= a.nq)
        return;
    uint8_t *stage = lds_rr + (uint64_t)wv * (RS_STAGE + 8ull * n2cap);
    uint64_t *kw = reinterpret_cast<uint64_t *>(stage + RS_STAGE);
    const uint8_t *qrow = a.queries + q * a.qstride;
    const uint32_t *cand = a.cand + q * a.cand_stride;
    const uint32_t m = frontier_len(a, q, lane), n2 = pow2_at_least(m);
    const uint32_t piece = lane % PPR, sub = lane / PPR;
    for (uint32_t g0 = 0; g0 < m; g0 += G) {
        const uint32_t gm = min(G, m - g0);
        const uint32_t my_id = lane < gm ? cand[g0 + lane] : 0u;
        // the next step's pieces are loaded into registers while this
        // step's are summed (two steps in flight per wave)
        rs_vec buf[8], qbuf = {0, 0, 0, 0};
        typename raw_acc<RT, IP>::T acc = 0;
        rs_fetch<PPR>(buf, qbuf, a.base, a.row_bytes, my_id, qrow, 0, gm, piece, sub);
        for (uint64_t c0 = 0; c0 < a.row_bytes; c0 += CH) {
            const bool in_row = c0 + 16u * piece < a.row_bytes;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if ((uint32_t)t * RPI + sub < gm && in_row)
                    *reinterpret_cast<rs_vec *>(stage + (t * RPI + sub) * PITCH + 16u * piece) = buf[t];
            if (sub == 0 && in_row)
                *reinterpret_cast<rs_vec *>(stage + G * PITCH + 16u * piece) = qbuf;
            if (c0 + CH < a.row_bytes)
                rs_fetch<PPR>(buf, qbuf, a.base, a.row_bytes, my_id, qrow, c0 + CH, gm, piece, sub);
            wave_sync();
            if (lane < gm)
                acc = raw_acc<RT, IP>::add(acc, stage + lane * PITCH, stage + G * PITCH,
                                           (uint32_t)min<uint64_t>(CH, a.row_bytes - c0));
            wave_sync();
        }
        if (lane < gm)
            kw[g0 + lane] = make_key(raw_acc<RT, IP>::fin(acc), my_id);
    }
    for (uint32_t i = m + lane; i < n2; i += WAVE)
        kw[i] = ~0ull;
    sort_and_write(a, q, kw, m, n2, lane);
}

template <typename F> static void rerank_dispatch(int raw_dtype, F &&go)
{
    using std::integral_constant;
    if (raw_dtype == NBPG_F32)
        go(integral_constant<int, NBPG_F32>());
    else if (raw_dtype == NBPG_U8)
        go(integral_constant<int, NBPG_U8>());
    else if (raw_dtype == NBPG_I8)
        go(integral_constant<int, NBPG_I8>());
    else
        go(integral_constant<int, NBPG_U4X2>()); // U4X2 and I4X2
}

int rerank_launch(int raw_dtype, uint32_t mips, const void *d_raw_base, uint64_t n, uint32_t dims,
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
    const bool staged = a.row_bytes % 16 == 0 && q_stride % 16 == 0 &&
                        ((reinterpret_cast<uintptr_t>(d_raw_base) |
                          reinterpret_cast<uintptr_t>(d_raw_queries)) & 15) == 0;
    if (staged) {
        uint32_t n2 = 1;
        while (n2 < nc)
            n2 <<= 1;
        // 4 waves per group while their LDS fits 64 KiB, else 2
        const uint32_t per_wave = RS_STAGE + 8 * n2, wpg = 4 * per_wave <= 65536 ? 4 : 2;
        const dim3 sg((uint32_t)((nq + wpg - 1) / wpg)), sb(WAVE * wpg);
        const uint32_t lds = wpg * per_wave;
        // few candidates: longer pieces of each row per step
        const int ppr = nc <= 16 ? 32 : (nc <= 32 ? 16 : 8);
        auto go = [&](auto RT) {
#define NBPG_RS_LAUNCH(P)                                                                          \
    do {                                                                                           \
        if (mips)                                                                                  \
            hipLaunchKernelGGL((rerank_staged_kernel<RT.value, true, P>), sg, sb, lds, s, a, n2);  \
        else                                                                                       \
            hipLaunchKernelGGL((rerank_staged_kernel<RT.value, false, P>), sg, sb, lds, s, a, n2); \
    } while (0)
            if (ppr == 32)
                NBPG_RS_LAUNCH(32);
            else if (ppr == 16)
                NBPG_RS_LAUNCH(16);
            else
                NBPG_RS_LAUNCH(8);
#undef NBPG_RS_LAUNCH
        };
        rerank_dispatch(raw_dtype, go);
    } else {
        auto go = [&](auto RT) {
            if (mips)
                hipLaunchKernelGGL((rerank_kernel<RT.value, true>), grid, block, 0, s, a);
            else
                hipLaunchKernelGGL((rerank_kernel<RT.value, false>), grid, block, 0, s, a);
        };
        rerank_dispatch(raw_dtype, go);
    }
    return hipGetLastError() == hipSuccess ? NBPG_OK : set_error(NBPG_EHIP, "rerank launch");
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_rerank(int raw_dtype, int metric, const void *d_raw_base, uint64_t n, uint32_t dims,
                           const void *d_raw_queries, uint64_t nq, const uint32_t *d_cand,
                           uint32_t cand_stride, uint32_t n_cand, uint32_t k,
                           uint32_t rerank_factor, uint32_t *d_out_ids, float *d_out_dists,
                           void *stream)
{
    NBPG_TRY(ensure_init());
    if (raw_dtype < NBPG_U8 || raw_dtype > NBPG_F32)
        return set_error(NBPG_EINVAL, "raw dtype %d", raw_dtype);
    if (metric != NBPG_METRIC_L2 && metric != NBPG_METRIC_MIPS)
        return set_error(NBPG_EINVAL, "metric %d", metric);
    if (nq && (!d_raw_base || !d_raw_queries || !d_cand || !d_out_ids))
        return set_error(NBPG_EINVAL, "null argument");
    if (k == 0 || rerank_factor == 0 || n_cand == 0 || cand_stride < n_cand || dims == 0)
        return set_error(NBPG_EINVAL, "k, rerank_factor and n_cand must be positive");
    return rerank_launch(raw_dtype, metric == NBPG_METRIC_MIPS ? 1u : 0u, d_raw_base, n, dims,
                         d_raw_queries,
                         (uint64_t)dims * raw_elem_bytes(raw_dtype), nq, d_cand, cand_stride,
                         n_cand, k, rerank_factor, d_out_ids, d_out_dists, as_stream(stream));
}
