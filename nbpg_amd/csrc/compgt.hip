// compgt.hip -- brute-force ground truth: the exact (dist, id) top-k per
// query, the result compute_groundtruth (CompGT.cpp:29-134) writes, with
// ties broken by id.
//
// Grid = query tiles x base slices.  A workgroup holds QT queries in LDS,
// streams its slice of the base (one row per thread per round), computes
// the QT exact int32 distances of each row, and keeps per query a sorted
// top-k list plus an admission buffer: a row enters the buffer only when
// its (dist, id) key beats the query's current k-th key, so after warm-up
// almost every distance is rejected by one compare.  Slices are merged by a
// second kernel.  Distances use |x|^2 + |q|^2 - 2<x,q> on the packed dot
// instructions (exact), or, in NBPG_GT_MFMA mode, int8 MFMA Gram tiles on
// the row-minus-128 codes (see compgt_mfma.hip).
#include <algorithm>

#include "internal.h"

namespace nbpg {

#define GT_QT 8
#define GT_THREADS 256
#define GT_KMAX 256

struct GtArgs {
    const uint8_t *base;
    const uint8_t *queries;
    uint64_t n, nq;
    uint32_t dims, k, slices;
    uint64_t slice_rows;
    uint64_t row_stride; // bytes between base rows (dims, or a multiple to sample)
    uint64_t *partial;   // [nq][slices][k] keys
    uint32_t mips;       // inner-product metric
    // NULL, or a device list of query numbers: list entry j is query
    // qidx[j], its partial results go to entry j, and the list's length is
    // *qcount, read on the device (so a launch can be sized for the most
    // entries and enqueued before the list is known)
    const uint32_t *qidx, *qcount;
};

// ascending bitonic sort of a[0:n2) by the whole workgroup
__device__ __forceinline__ void block_bitonic(uint64_t *a, uint32_t n2)
{
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = a[i], y = a[p];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ascending bitonic sort of a[0:n2) by one wavefront (n2 <= 256 here)
__device__ __forceinline__ void gt_wave_bitonic(uint64_t *a, uint32_t n2, uint32_t lane)
{
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n2; i += WAVE) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = a[i], y = a[p];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[p] = x;
                    }
                }
            }
            wave_sync();
        }
    }
}

// merge buf[0:nb) (unsorted) into the sorted top[0:tn); one wavefront
__device__ void gt_merge(uint64_t *top, uint32_t &tn, uint64_t *buf, uint32_t nb, uint64_t *tmp,
                         uint32_t k, uint32_t lane)
{
    uint32_t n2 = 1;
    while (n2 < nb)
        n2 <<= 1;
    for (uint32_t i = nb + lane; i < n2; i += WAVE)
        buf[i] = ~0ull;
    wave_sync();
    gt_wave_bitonic(buf, n2, lane);
    for (uint32_t i = lane; i < tn; i += WAVE) {
        const uint64_t v = top[i];
        const uint32_t p = i + count_less(buf, nb, v);
        if (p < k)
            tmp[p] = v;
    }
    for (uint32_t j = lane; j < nb; j += WAVE) {
        const uint64_t v = buf[j];
        const uint32_t p = j + count_less(top, tn, v);
        if (p < k)
            tmp[p] = v;
    }
    const uint32_t nt = min(k, tn + nb);
    wave_sync();
    for (uint32_t i = lane; i < nt; i += WAVE)
        top[i] = tmp[i];
    tn = nt;
    wave_sync();
}

template <int DT, int NCH>
__global__ __launch_bounds__(GT_THREADS) void compgt_plain_kernel(GtArgs a)
{
    constexpr int NW = NCH * 4;
    __shared__ uint32_t qs[GT_QT][NW];
    __shared__ uint32_t qn[GT_QT];
    __shared__ uint64_t top[GT_QT][GT_KMAX];
    __shared__ uint64_t buf[GT_QT][GT_THREADS];
    __shared__ uint64_t tmp[GT_THREADS / WAVE][GT_KMAX];
    __shared__ uint32_t cnt[GT_QT], tnum[GT_QT];
    __shared__ uint64_t thr[GT_QT];
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    const uint64_t nq = a.qidx ? *a.qcount : a.nq;
    const uint32_t words = a.dims / 4;
    // query tiles: one per workgroup row, or a loop over a device list
    for (uint64_t tile = blockIdx.x; tile * GT_QT < nq; tile += gridDim.x) {
    const uint64_t q0 = tile * GT_QT;
    const uint32_t nqt = (uint32_t)min<uint64_t>(GT_QT, nq - q0);
    for (uint32_t i = tid; i < GT_QT * NW; i += GT_THREADS) {
        const uint32_t q = i / NW, w = i % NW;
        uint32_t v = 0;
        if (q < nqt && w < words) {
            const uint64_t qr = a.qidx ? a.qidx[q0 + q] : q0 + q;
            v = reinterpret_cast<const uint32_t *>(a.queries + qr * a.dims)[w];
        }
        qs[q][w] = v;
    }
    if (tid < GT_QT) {
        cnt[tid] = 0;
        tnum[tid] = 0;
        thr[tid] = ~0ull;
    }
    __syncthreads();
    if (tid < GT_QT) {
        uint32_t acc = 0;
        for (int w = 0; w < NW; ++w)
            acc = dot_traits<DT>::dot(qs[tid][w], qs[tid][w], acc);
        qn[tid] = acc;
    }
    const uint64_t r0 = (uint64_t)blockIdx.y * a.slice_rows;
    const uint64_t r1 = min<uint64_t>(a.n, r0 + a.slice_rows);
    __syncthreads();
    for (uint64_t rb = r0; rb < r1; rb += GT_THREADS) {
        const uint64_t r = rb + tid;
        if (r < r1) {
            uint32_t x[NW];
            const uint32_t *rw = reinterpret_cast<const uint32_t *>(a.base + r * a.row_stride);
#pragma unroll
            for (int w = 0; w < NW; ++w)
                x[w] = (uint32_t)w < words ? rw[w] : 0u;
            uint32_t nx = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w)
                nx = dot_traits<DT>::dot(x[w], x[w], nx);
            for (uint32_t q = 0; q < nqt; ++q) {
                uint32_t acc = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w)
                    acc = dot_traits<DT>::dot(x[w], qs[q][w], acc);
                const uint64_t key = make_key(dist_from<DT>(nx, qn[q], acc, a.mips), (uint32_t)r);
                if (key < thr[q]) {
                    const uint32_t p = atomicAdd(&cnt[q], 1u);
                    buf[q][p] = key;
                }
            }
        }
        __syncthreads();
        for (uint32_t q = wv; q < nqt; q += GT_THREADS / WAVE) {
            const uint32_t nb = cnt[q];
            if (nb) {
                uint32_t tn = tnum[q];
                gt_merge(top[q], tn, buf[q], nb, tmp[wv], a.k, lane);
                if (lane == 0) {
                    tnum[q] = tn;
                    cnt[q] = 0;
                    thr[q] = tn == a.k ? top[q][a.k - 1] : ~0ull;
                }
            }
        }
        __syncthreads();
    }
    for (uint32_t q = 0; q < nqt; ++q)
        for (uint32_t i = tid; i < a.k; i += GT_THREADS)
            a.partial[((q0 + q) * a.slices + blockIdx.y) * a.k + i] = i < tnum[q] ? top[q][i] : ~0ull;
    __syncthreads(); // the next tile reuses the shared lists
    }
}

// per query: sort the slices' candidates, emit k
// qidx / qcount as in GtArgs: entry j of the list is written to query qidx[j]
__global__ __launch_bounds__(256) void compgt_merge_kernel(const uint64_t *__restrict__ partial,
                                                          uint32_t slices, uint32_t k,
                                                          uint32_t *__restrict__ ids,
                                                          float *__restrict__ dists,
                                                          uint64_t nq_launch,
                                                          const uint32_t *__restrict__ qidx,
                                                          const uint32_t *__restrict__ qcount)
{
    extern __shared__ uint64_t all[];
    const uint64_t nq = qidx ? *qcount : nq_launch;
    const uint32_t m = slices * k;
    uint32_t n2 = 1;
    while (n2 < m)
        n2 <<= 1;
    for (uint64_t q = blockIdx.x; q < nq; q += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x)
            all[i] = i < m ? partial[q * m + i] : ~0ull;
        __syncthreads();
        block_bitonic(all, n2);
        const uint64_t o = qidx ? qidx[q] : q;
        for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
            ids[o * k + i] = key_id(all[i]);
            if (dists)
                dists[o * k + i] = key_dist(all[i]);
        }
        __syncthreads();
    }
}

int compgt_merge_launch(const uint64_t *partial, uint64_t nq, uint32_t slices, uint32_t k,
                        uint32_t *ids, float *dists, hipStream_t s, const uint32_t *qidx,
                        const uint32_t *qcount)
{
    uint32_t m = slices * k, n2 = 1;
    while (n2 < m)
        n2 <<= 1;
    const uint32_t grid = (uint32_t)(qidx ? std::min<uint64_t>(nq, 2048) : nq);
    hipLaunchKernelGGL(compgt_merge_kernel, dim3(grid), dim3(256), n2 * 8, s, partial, slices, k,
                       ids, dists, nq, qidx, qcount);
    return hipGetLastError() == hipSuccess ? NBPG_OK : set_error(NBPG_EHIP, "compgt merge launch");
}

int compgt_mfma(int dtype, const uint8_t *base, uint64_t n, const uint8_t *queries, uint64_t nq,
                uint32_t dims, uint32_t k, uint32_t mips, uint32_t *out_ids, float *out_dists,
                hipStream_t stream);

// exact top-k from distance tiles; base row r at base + r * row_stride
int compgt_tiles(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                 const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k, bool use_mfma,
                 uint32_t mips, uint32_t *ids, float *dists, hipStream_t s);

int compgt_refrule(int dtype, const uint8_t *base, uint64_t n, const uint8_t *queries, uint64_t nq,
                   uint32_t dims, uint32_t k, bool use_mfma, uint32_t mips, uint32_t *ids,
                   float *dists, hipStream_t s);

// slices of the base per query tile for `tiles` tiles: enough workgroups to
// fill the device, slices of at least 2048 rows, at most 4096 / k of them
static uint64_t plain_slices(uint64_t tiles, uint64_t n, uint32_t k)
{
    uint64_t slices = std::max<uint64_t>(1, (uint64_t)g_dev.cu * 8 / std::max<uint64_t>(1, tiles));
    slices = std::min<uint64_t>(slices, std::max<uint64_t>(1, n / 2048));
    return std::max<uint64_t>(1, std::min<uint64_t>(slices, 4096 / k));
}

// bytes of the partial buffer compgt_plain_core needs for nq queries
uint64_t compgt_plain_partial_bytes(uint64_t nq, uint64_t n, uint32_t k, bool listed)
{
    const uint64_t tiles = (nq + GT_QT - 1) / GT_QT;
    return nq * plain_slices(listed ? std::min<uint64_t>(tiles, 256) : tiles, n, k) * k * 8;
}

// plain scan + slice merge into ids / dists (dists may be NULL), with a
// caller-provided partial buffer (compgt_plain_partial_bytes).  Base row r
// is at base + r * row_stride and is reported as id r.  qidx / qcount: an
// optional device list of at most nq query numbers (see GtArgs); the launch
// is then a loop over at most 256 query tiles' worth of workgroups.
int compgt_plain_core(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                      const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k,
                      uint32_t mips, uint32_t *ids, float *dists, hipStream_t s,
                      uint64_t *partial, const uint32_t *qidx, const uint32_t *qcount)
{
    const uint32_t nch = nch_for_bytes(dims);
    uint64_t tiles = (nq + GT_QT - 1) / GT_QT;
    if (qidx)
        tiles = std::min<uint64_t>(tiles, 256);
    const uint64_t slices = plain_slices(tiles, n, k);
    const uint64_t slice_rows = (n + slices - 1) / slices;
    GtArgs a;
    a.base = base;
    a.queries = queries;
    a.n = n;
    a.nq = nq;
    a.dims = dims;
    a.k = k;
    a.slices = (uint32_t)slices;
    a.slice_rows = slice_rows;
    a.row_stride = row_stride;
    a.partial = partial;
    a.mips = mips;
    a.qidx = qidx;
    a.qcount = qcount;
    int rc = dispatch_dt_nch(dtype, nch, [&](auto DT, auto NCH) {
        hipLaunchKernelGGL((compgt_plain_kernel<DT.value, NCH.value>),
                           dim3((uint32_t)tiles, (uint32_t)slices), dim3(GT_THREADS), 0, s, a);
    });
    if (rc == NBPG_OK && hipGetLastError() != hipSuccess)
        rc = set_error(NBPG_EHIP, "compgt launch");
    if (rc == NBPG_OK)
        rc = compgt_merge_launch(partial, nq, (uint32_t)slices, k, ids, dists, s, qidx, qcount);
    return rc;
}

int compgt_plain_launch(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                        const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k,
                        uint32_t mips, uint32_t *ids, float *dists, hipStream_t s)
{
    uint64_t *partial = nullptr;
    if (hipMallocAsync((void **)&partial, compgt_plain_partial_bytes(nq, n, k, false), s) !=
        hipSuccess)
        return set_error(NBPG_ENOMEM, "compgt partial buffer");
    const int rc = compgt_plain_core(dtype, base, n, row_stride, queries, nq, dims, k, mips, ids,
                                     dists, s, partial, nullptr, nullptr);
    hipFreeAsync(partial, s);
    return rc;
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_compgt(int dtype, const void *d_base, uint64_t n, const void *d_queries,
                           uint64_t nq, uint32_t dims, uint32_t k, int mode, uint32_t *d_out_ids,
                           float *d_out_dists, void *stream)
{
    NBPG_TRY(ensure_init());
    if ((nq && (!d_base || !d_queries || !d_out_ids)) || dtype < NBPG_U8 || dtype > NBPG_I4X2)
        return set_error(NBPG_EINVAL, "bad argument");
    if (k == 0 || k > GT_KMAX || n < k) // CompGT.cpp:120-123 reads k entries
        return set_error(NBPG_EINVAL, "k = %u with n = %llu (1 <= k <= min(n, %d))", k,
                         (unsigned long long)n, GT_KMAX);
    if (dims % 4 != 0 || (reinterpret_cast<uintptr_t>(d_base) & 3) ||
        (reinterpret_cast<uintptr_t>(d_queries) & 3))
        return set_error(NBPG_EUNSUP, "rows must be whole 4-byte words (dims %u)", dims);
    if (nch_for_bytes(dims) == 0)
        return set_error(NBPG_EUNSUP, "rows of %u bytes (<= 256 supported)", dims);
    if (n >= 0xFFFFFFFFull)
        return set_error(NBPG_EUNSUP, "n too large");
    if (nq == 0)
        return NBPG_OK;
    hipStream_t s = as_stream(stream);
    const uint32_t mips = (mode & NBPG_GT_MIPS) ? 1u : 0u;
    mode &= ~NBPG_GT_MIPS;
    if (mode & NBPG_GT_REFTIES) {
        const int engine = mode & ~NBPG_GT_REFTIES;
        if (engine == NBPG_GT_MFMA &&
            ((dtype != NBPG_U8 && dtype != NBPG_I8) || dims % 16 ||
             ((reinterpret_cast<uintptr_t>(d_base) | reinterpret_cast<uintptr_t>(d_queries)) & 15)))
            return set_error(NBPG_EUNSUP, "MFMA ground truth needs 16-byte u8 / i8 rows");
        if (engine != NBPG_GT_MFMA && engine != NBPG_GT_PLAIN)
            return set_error(NBPG_EINVAL, "compgt mode %d", mode);
        return compgt_refrule(dtype, (const uint8_t *)d_base, n, (const uint8_t *)d_queries, nq,
                              dims, k, engine == NBPG_GT_MFMA, mips, d_out_ids, d_out_dists, s);
    }
    if (mode == NBPG_GT_MFMA) {
        if ((dtype != NBPG_U8 && dtype != NBPG_I8) || dims % 16 ||
            ((reinterpret_cast<uintptr_t>(d_base) | reinterpret_cast<uintptr_t>(d_queries)) & 15))
            return set_error(NBPG_EUNSUP, "MFMA ground truth needs 16-byte u8 / i8 rows");
        return compgt_mfma(dtype, (const uint8_t *)d_base, n, (const uint8_t *)d_queries, nq, dims,
                           k, mips, d_out_ids, d_out_dists, s);
    }
    if (mode == NBPG_GT_PLAIN)
        return compgt_plain_launch(dtype, (const uint8_t *)d_base, n, dims,
                                   (const uint8_t *)d_queries, nq, dims, k, mips, d_out_ids,
                                   d_out_dists, s);
    return set_error(NBPG_EINVAL, "compgt mode %d", mode);
}
