// start.hip -- per-query start points for the FAST search: each query
// starts from the nearest of the index's seed rows.
//
// The reference's driver asks get_sp(i) for query i's start points
// (beamSearch.h:257-303 with check_nn_recall.h:282-297); here the caller
// gives the index a set of up to NBPG_MAX_SEEDS seed rows once
// (nbpg_index_set_seeds), and a pass before the search finds, for every
// query, the seed at the smallest exact distance (ties to the lower id).
// Starting next to the query shortens its walk through the graph: fewer
// expansions and fewer row gathers for the same recall.
//
// The pass is a queries x seeds product on int8 matrix cores
// (v_mfma_i32_16x16x64_i8): seed rows are kept as signed bytes x' (u8 rows
// as x - 128, i8 rows as they are; bytes past dims zero), so that per seed s
//     squared L2:    |s - q|^2 = |s'|^2 - 2 <s', q'> + |q'|^2
//     inner product: -<s, q>   = -(<s', q'> + A(s) + A(q)),
//                    A(x) = 128 sum(x') + 8192 dims (u8), 0 (i8)
// and the query's own term is the same for every seed: the argmin is that
// of e(s) = c(s) - g <s', q'> with c(s) = |s'|^2, g = 2 (L2) or c(s) = -A(s),
// g = 1 (inner product), an exact int32.  One workgroup of 8 wavefronts
// takes 32 queries (two 16-query tiles held as MFMA operands per
// wavefront; 313 workgroups for 10k queries, so every CU takes part); the
// wavefronts split the seed tiles between them, each lane keeps the smallest
// e of the seeds it sees (the first on ties: its seeds come in ascending
// position), and the workgroup combines the (e, seed position) minima in
// LDS.  The search
// kernel reads the chosen id from the query's first output slot, which it
// overwrites with its results only at the end.
#include <algorithm>
#include <climits>
#include <vector>

#include "search_args.h"

namespace nbpg {

typedef int st_v4i __attribute__((ext_vector_type(4)));

#define ST_WAVES 8 // wavefronts per workgroup
#define ST_QT 2    // 16-query tiles per wavefront: 32 queries per workgroup
#define ST_PAD_SELF 0x3FFFFFFF // seed term of a padding row: never the minimum

// seed i of the padded tile set: its row as signed bytes (zero past dims and
// for padding rows) and its seed terms for both metrics
__global__ void seed_rows_kernel(const uint8_t *__restrict__ vecs, uint32_t stride, uint32_t dims,
                                 int is_u8, const uint32_t *__restrict__ ids, uint32_t count,
                                 uint32_t padded, uint8_t *__restrict__ rows,
                                 int32_t *__restrict__ self_l2, int32_t *__restrict__ self_ip)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= padded)
        return;
    int32_t sq = 0, sum = 0;
    const uint8_t *src = i < count ? vecs + (uint64_t)ids[i] * stride : nullptr;
    for (uint32_t b = 0; b < stride; ++b) {
        int8_t v = 0;
        if (src && b < dims)
            v = (int8_t)(is_u8 ? (src[b] ^ 0x80u) : src[b]);
        rows[(uint64_t)i * stride + b] = (uint8_t)v;
        sq += (int32_t)v * v;
        sum += v;
    }
    self_l2[i] = i < count ? sq : ST_PAD_SELF;
    self_ip[i] = i < count ? (is_u8 ? -(128 * sum + 8192 * (int32_t)dims) : 0) : ST_PAD_SELF;
}

struct StartArgs {
    const uint8_t *queries;
    uint64_t qstride, nq;
    uint32_t dims, stride; // query bytes used; seed row pitch
    int is_u8;
    uint32_t tiles;        // 16-seed tiles
    const uint8_t *rows;   // tiles * 16 seed rows as signed bytes
    const int32_t *self;   // seed term per seed row (the metric's)
    const uint32_t *ids;   // seed ids, ascending
    uint32_t shift;        // g = 1 << shift
    uint32_t *out;         // query q's start id -> out[q * out_k]
    uint32_t out_k;
};

template <int KS> __global__ __launch_bounds__(64 * ST_WAVES) void start_select_kernel(StartArgs a)
{
    // seed tiles whose loads a wavefront issues together (fewer for wide
    // rows, whose fragments take more registers)
    constexpr int ST_GROUP = KS <= 2 ? 4 : (KS == 3 ? 2 : 1);
    __shared__ uint64_t best[ST_WAVES][16 * ST_QT];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint32_t col = lane % 16, kq = lane / 16;
    const uint64_t q0 = (uint64_t)blockIdx.x * (16 * ST_QT);
    const uint32_t xm = a.is_u8 ? 0x80808080u : 0u;
    const int32_t gneg = -(1 << a.shift);
    // query tiles as B operands: query q0 + 16 t + col, bytes 64 s + 16 kq
    st_v4i qf[ST_QT][KS];
#pragma unroll
    for (int t = 0; t < ST_QT; ++t) {
        const uint64_t q = q0 + 16 * t + col;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint32_t off = 64 * s + 16 * kq;
            st_v4i v = {0, 0, 0, 0};
            // dims is a multiple of 16: a 16-byte piece is all in or all out
            if (q < a.nq && off < a.dims) {
                const uint4 u = *reinterpret_cast<const uint4 *>(a.queries + q * a.qstride + off);
                v = st_v4i{(int)(u.x ^ xm), (int)(u.y ^ xm), (int)(u.z ^ xm), (int)(u.w ^ xm)};
            }
            qf[t][s] = v;
        }
    }
    // per query tile: the smallest e this lane has seen and its seed position
    int32_t be[ST_QT];
    uint32_t bp[ST_QT];
#pragma unroll
    for (int t = 0; t < ST_QT; ++t) {
        be[t] = INT_MAX;
        bp[t] = 0xFFFFFFFFu;
    }
    // lane l holds D[4 kq + r][col]: seed 16 st + 4 kq + r against query col.
    // The wavefront's seed tiles go ST_GROUP at a time, every load of a
    // group issued before its first product (the tiles come from L2, and a
    // wavefront has only a few of them: latency, not bandwidth, decides)
    for (uint32_t g0 = wv; g0 < a.tiles; g0 += ST_WAVES * ST_GROUP) {
        st_v4i sf[ST_GROUP][KS], c4[ST_GROUP];
#pragma unroll
        for (int j = 0; j < ST_GROUP; ++j) {
            const uint32_t st = g0 + ST_WAVES * j;
            const bool ok = st < a.tiles;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint32_t off = 64 * s + 16 * kq;
                sf[j][s] = ok && off < a.stride
                               ? *reinterpret_cast<const st_v4i *>(
                                     a.rows + (uint64_t)(16 * st + col) * a.stride + off)
                               : st_v4i{0, 0, 0, 0};
            }
            c4[j] = ok ? *reinterpret_cast<const st_v4i *>(a.self + 16 * st + 4 * kq)
                       : st_v4i{ST_PAD_SELF, ST_PAD_SELF, ST_PAD_SELF, ST_PAD_SELF};
        }
#pragma unroll
        for (int j = 0; j < ST_GROUP; ++j) {
            const uint32_t st = g0 + ST_WAVES * j;
#pragma unroll
            for (int t = 0; t < ST_QT; ++t) {
                st_v4i acc = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS; ++s)
                    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(sf[j][s], qf[t][s], acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    // |acc| < 2^23: one 24-bit multiply-add
                    const int32_t e = __mul24(acc[r], gneg) + c4[j][r];
                    const bool take = e < be[t];
                    be[t] = take ? e : be[t];
                    bp[t] = take ? 16 * st + 4 * kq + r : bp[t];
                }
            }
        }
    }
    // the four lanes of a query column, then the workgroup's wavefronts
#pragma unroll
    for (int t = 0; t < ST_QT; ++t) {
        uint64_t k = ((uint64_t)((uint32_t)be[t] ^ 0x80000000u) << 32) | bp[t];
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)k, o);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(k >> 32), o);
            const uint64_t v = ((uint64_t)hi << 32) | lo;
            k = v < k ? v : k;
        }
        if (kq == 0)
            best[wv][16 * t + col] = k;
    }
    __syncthreads();
    if (threadIdx.x < 16 * ST_QT) {
        const uint64_t q = q0 + threadIdx.x;
        if (q < a.nq) {
            uint64_t m = best[0][threadIdx.x];
            for (int w = 1; w < ST_WAVES; ++w)
                m = best[w][threadIdx.x] < m ? best[w][threadIdx.x] : m;
            a.out[q * a.out_k] = a.ids[(uint32_t)m];
        }
    }
}

bool start_select_supported(const nbpg_index *ix)
{
    return (ix->dtype == NBPG_U8 || ix->dtype == NBPG_I8) && ix->dims % 16 == 0 &&
           ix->stride <= 256;
}

int start_select(const nbpg_index *ix, const void *d_queries, uint64_t nq, uint64_t qstride,
                 uint32_t *d_out_ids, uint32_t out_k, hipStream_t stream)
{
    if (ix->n_seeds == 0)
        return set_error(NBPG_EINVAL, "no seed rows on the index (nbpg_index_set_seeds)");
    if (!start_select_supported(ix))
        return set_error(NBPG_EUNSUP, "nearest-seed starts need u8 / i8 rows of a multiple of "
                                      "16 bytes (dims %u)", ix->dims);
    if ((qstride & 15) || (reinterpret_cast<uintptr_t>(d_queries) & 15))
        return set_error(NBPG_EINVAL, "nearest-seed starts need 16-byte aligned query rows");
    if (!d_out_ids)
        return set_error(NBPG_EINVAL, "nearest-seed starts need the id output");
    StartArgs a;
    memset(&a, 0, sizeof(a));
    a.queries = (const uint8_t *)d_queries;
    a.qstride = qstride;
    a.nq = nq;
    a.dims = ix->dims;
    a.stride = ix->stride;
    a.is_u8 = ix->dtype == NBPG_U8;
    a.tiles = (ix->n_seeds + 15) / 16;
    a.rows = ix->seed_rows;
    a.self = ix->mips ? ix->seed_self_ip : ix->seed_self_l2;
    a.ids = ix->seed_ids;
    a.shift = ix->mips ? 0u : 1u;
    a.out = d_out_ids;
    a.out_k = out_k;
    const dim3 grid((uint32_t)((nq + 16 * ST_QT - 1) / (16 * ST_QT))), block(64 * ST_WAVES);
    switch ((ix->stride + 63) / 64) {
    case 1: hipLaunchKernelGGL(start_select_kernel<1>, grid, block, 0, stream, a); break;
    case 2: hipLaunchKernelGGL(start_select_kernel<2>, grid, block, 0, stream, a); break;
    case 3: hipLaunchKernelGGL(start_select_kernel<3>, grid, block, 0, stream, a); break;
    default: hipLaunchKernelGGL(start_select_kernel<4>, grid, block, 0, stream, a); break;
    }
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

void drop_seeds(nbpg_index *ix)
{
    if (ix->seed_rows)
        hipDeviceSynchronize(); // a search may still read them
    hipFree(ix->seed_rows);
    hipFree(ix->seed_self_l2);
    hipFree(ix->seed_self_ip);
    hipFree(ix->seed_ids);
    ix->seed_rows = nullptr;
    ix->seed_self_l2 = ix->seed_self_ip = nullptr;
    ix->seed_ids = nullptr;
    ix->n_seeds = 0;
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_index_set_seeds(nbpg_index *ix, const uint32_t *h_ids, uint32_t count)
{
    NBPG_TRY(ensure_init());
    if (!ix || (count && !h_ids))
        return set_error(NBPG_EINVAL, "null argument");
    if (count > NBPG_MAX_SEEDS)
        return set_error(NBPG_EINVAL, "%u seeds (at most %d)", count, NBPG_MAX_SEEDS);
    std::vector<uint32_t> ids(h_ids, h_ids + count);
    std::sort(ids.begin(), ids.end());
    for (uint32_t i = 0; i < count; ++i) {
        if (ids[i] >= ix->n)
            return set_error(NBPG_EINVAL, "seed %u >= n", ids[i]);
        if (i && ids[i] == ids[i - 1])
            return set_error(NBPG_EINVAL, "seed %u repeated", ids[i]);
    }
    drop_seeds(ix);
    if (count == 0)
        return NBPG_OK;
    if (!start_select_supported(ix))
        return set_error(NBPG_EUNSUP, "seed rows need u8 / i8 rows of a multiple of 16 bytes");
    const uint32_t padded = (count + 15) / 16 * 16;
    const bool ok = hipMalloc(&ix->seed_rows, (uint64_t)padded * ix->stride) == hipSuccess &&
                    hipMalloc(&ix->seed_self_l2, padded * 4ull) == hipSuccess &&
                    hipMalloc(&ix->seed_self_ip, padded * 4ull) == hipSuccess &&
                    hipMalloc(&ix->seed_ids, padded * 4ull) == hipSuccess;
    if (!ok) {
        drop_seeds(ix);
        return set_error(NBPG_ENOMEM, "seed rows");
    }
    ids.resize(padded, 0u); // padding entries are never chosen (seed term ST_PAD_SELF)
    HIP_TRY(hipMemcpy(ix->seed_ids, ids.data(), padded * 4ull, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(seed_rows_kernel, dim3((padded + 255) / 256), dim3(256), 0, 0, ix->vecs,
                       ix->stride, ix->dims, ix->dtype == NBPG_U8 ? 1 : 0, ix->seed_ids, count,
                       padded, ix->seed_rows, ix->seed_self_l2, ix->seed_self_ip);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    ix->n_seeds = count;
    return NBPG_OK;
}
