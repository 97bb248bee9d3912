// compgt_ref.hip -- ground truth under the reference's own admission rule.
//
// compute_groundtruth (CompGT.cpp:63-97) keeps k array slots per query.  The
// first k rows fill them while the first slot holding the largest distance is
// tracked (strict >).  A later row with dist < topdist replaces slot toppos;
// then topdist / toppos become the first maximum over the slots, where the
// scan starts from the value slot 0 held BEFORE the replacement
// (CompGT.cpp:82-84 reads topk[0] ahead of the write).  When the maximum sat
// in slot 0 the recomputed topdist is therefore stale, the next admission
// evicts slot 0 whatever it holds, and the result is not always the exact
// top-k; with ties at the k-th distance the surviving ids depend on slot
// positions.  This kernel replays that sequence row by row in id order and
// sorts the k slots by (dist, id) at the end (CompGT.cpp:102-118).
//
// The same distance tiles also drive the exact (dist, id) top-k
// (compgt_tile_topk_kernel): every query keeps a sorted list of its k
// smallest keys in LDS and a row enters it only when its distance is at most
// the list's k-th (one float compare and a ballot per 64 rows); the base is
// cut into slices whose lists are merged afterwards.
//
// A workgroup owns RQ queries and walks the whole base in 256-row chunks:
// its four wavefronts first compute the RQ x 256 distance tile (int8 MFMA
// Gram tiles for 16-byte u8 / i8 rows, packed dot instructions otherwise),
// then each wavefront replays the rule for RQ/4 queries.  The replay jumps
// from admission to admission with a ballot of `dist < topdist` over 64 rows
// and recomputes topdist / toppos with one wave reduction over the k slots.
#include <algorithm>

#include "internal.h"

namespace nbpg {

// compgt.hip: sorts each query's slices x k keys and writes the first k
int compgt_merge_launch(const uint64_t *partial, uint64_t nq, uint32_t slices, uint32_t k,
                        uint32_t *ids, float *dists, hipStream_t s,
                        const uint32_t *qidx = nullptr, const uint32_t *qcount = nullptr);

typedef int v4i __attribute__((ext_vector_type(4)));

#define RG_CH 256              // rows per chunk
#define RG_TS (RG_CH + 4)      // distance tile row stride (floats)
#define RG_THREADS 256

struct RefGtArgs {
    const uint8_t *base;
    const uint8_t *queries;
    uint64_t n, nq;
    uint32_t dims, k;
    uint32_t *ids;
    float *dists;
    uint32_t mips;       // inner-product metric
    uint64_t row_stride; // bytes between base rows (dims, or a multiple to sample)
};

__device__ __forceinline__ uint64_t wave_max64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, o), hi = __shfl_xor((uint32_t)(v >> 32), o);
        const uint64_t w = ((uint64_t)hi << 32) | lo;
        v = w > v ? w : v;
    }
    return v;
}

// (first maximum) over slots 0..k-1 where slot 0 reads as v0: the key
// orders by value, then by lower slot index
__device__ __forceinline__ uint64_t first_max(const float *sd, uint32_t k, float v0, uint32_t lane)
{
    uint64_t best = 0;
    for (uint32_t l = lane; l < k; l += WAVE) {
        const float v = l == 0 ? v0 : sd[l];
        const uint64_t key = ((uint64_t)ord_f(v) << 32) | (~l & 0xFFFFFFFFu);
        best = key > best ? key : best;
    }
    return wave_max64(best);
}

__device__ __forceinline__ void rg_wave_bitonic(uint64_t *a, uint32_t n2, uint32_t lane)
{
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n2; i += WAVE) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = a[i], y = a[p];
                    if ((x > y) == ((i & kk) == 0)) {
                        a[i] = y;
                        a[p] = x;
                    }
                }
            }
            wave_sync();
        }
}

// distance tile producers: D[q][c] = dist(query q0+q, row c0+c), +inf past n
template <int KS, int QT> struct MfmaTile {
    static constexpr int RQ = 16 * QT;
    v4i a[QT][KS];
    uint32_t qn[QT][4]; // |q'|^2, or A(q) under the inner product (compgt_mfma.hip)
    uint32_t xmask, is_u8_, mips;

    // |v|^2 of the row's words, or sum(v) for the inner product
    __device__ int part_of(const v4i &v, int part) const
    {
        const int o = 0x01010101;
        part = __builtin_amdgcn_sdot4(v.x, mips ? o : v.x, part, false);
        part = __builtin_amdgcn_sdot4(v.y, mips ? o : v.y, part, false);
        part = __builtin_amdgcn_sdot4(v.z, mips ? o : v.z, part, false);
        part = __builtin_amdgcn_sdot4(v.w, mips ? o : v.w, part, false);
        return part;
    }
    // <x,q> = <x',q'> + A(x) + A(q), A = 128 sum(x') + 8192 dims for u8, 0 for i8
    __device__ int a_of(int sum, uint32_t dims) const
    {
        return is_u8_ ? 128 * sum + 8192 * (int)dims : 0;
    }

    __device__ void init(const RefGtArgs &g, uint64_t q0, uint32_t is_u8, void *, uint32_t lane)
    {
        xmask = is_u8 ? 0x80808080u : 0u;
        is_u8_ = is_u8;
        mips = g.mips;
        const uint32_t row16 = lane % 16, kq = lane / 16;
#pragma unroll
        for (int t = 0; t < QT; ++t) {
            const uint64_t q = q0 + 16 * t + row16;
            int part = 0;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint32_t off = 64 * s + 16 * kq;
                v4i v = {0, 0, 0, 0};
                if (q < g.nq && off < g.dims) {
                    const uint4 u = *reinterpret_cast<const uint4 *>(g.queries + q * g.dims + off);
                    v = v4i{(int)(u.x ^ xmask), (int)(u.y ^ xmask), (int)(u.z ^ xmask),
                            (int)(u.w ^ xmask)};
                }
                a[t][s] = v;
                part = part_of(v, part);
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32); // |q'|^2 (or sum(q')) of query 16t + row16
            if (mips)
                part = a_of(part, g.dims);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                qn[t][r] = (uint32_t)__shfl(part, 4 * kq + r);
        }
    }

    // wave wv fills columns [64 wv, 64 wv + 64)
    __device__ void fill(const RefGtArgs &g, uint64_t c0, float *D, uint32_t wv, uint32_t lane)
    {
        const uint32_t row16 = lane % 16, kq = lane / 16;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t col = 64 * wv + 16 * b + row16;
            const uint64_t row = c0 + col;
            const bool rv = row < g.n;
            v4i x[KS];
            int part = 0;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint32_t off = 64 * s + 16 * kq;
                v4i v = {0, 0, 0, 0};
                if (rv && off < g.dims) {
                    const uint4 u = *reinterpret_cast<const uint4 *>(g.base + row * g.row_stride + off);
                    v = v4i{(int)(u.x ^ xmask), (int)(u.y ^ xmask), (int)(u.z ^ xmask),
                            (int)(u.w ^ xmask)};
                }
                x[s] = v;
                part = part_of(v, part);
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32); // |x'|^2 (or sum(x')) of this lane's row
            if (mips)
                part = a_of(part, g.dims);
#pragma unroll
            for (int t = 0; t < QT; ++t) {
                v4i c = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS; ++s)
                    c = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[t][s], x[s], c, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float d =
                        mips ? -(float)(int32_t)((uint32_t)c[r] + qn[t][r] + (uint32_t)part)
                             : (float)(int32_t)(qn[t][r] + (uint32_t)part - 2u * (uint32_t)c[r]);
                    D[(16 * t + 4 * kq + r) * RG_TS + col] = rv ? d : __builtin_huge_valf();
                }
            }
        }
    }
};

template <int DT, int NCH, int QT> struct DotTile {
    static constexpr int RQ = 16 * QT;
    static constexpr int NW = 4 * NCH;
    uint32_t *qs; // [RQ][NW] words in LDS
    uint32_t *qn; // [RQ]

    __device__ void init(const RefGtArgs &g, uint64_t q0, uint32_t, void *lds, uint32_t)
    {
        qs = reinterpret_cast<uint32_t *>(lds);
        qn = qs + RQ * NW;
        const uint32_t words = g.dims / 4;
        for (uint32_t i = threadIdx.x; i < RQ * NW; i += blockDim.x) {
            const uint32_t q = i / NW, w = i % NW;
            uint32_t v = 0;
            if (q0 + q < g.nq && w < words)
                v = reinterpret_cast<const uint32_t *>(g.queries + (q0 + q) * g.dims)[w];
            qs[i] = v;
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < RQ; q += blockDim.x) {
            uint32_t acc = 0;
            for (int w = 0; w < NW; ++w)
                acc = dot_traits<DT>::dot(qs[q * NW + w], qs[q * NW + w], acc);
            qn[q] = acc;
        }
        __syncthreads();
    }

    // thread t fills column t
    __device__ void fill(const RefGtArgs &g, uint64_t c0, float *D, uint32_t, uint32_t)
    {
        const uint32_t col = threadIdx.x;
        const uint64_t row = c0 + col;
        const bool rv = row < g.n;
        const uint32_t words = g.dims / 4;
        uint32_t x[NW];
        const uint32_t *rw = reinterpret_cast<const uint32_t *>(g.base + (rv ? row : 0) * g.row_stride);
#pragma unroll
        for (int w = 0; w < NW; ++w)
            x[w] = (rv && (uint32_t)w < words) ? rw[w] : 0u;
        uint32_t nx = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w)
            nx = dot_traits<DT>::dot(x[w], x[w], nx);
        for (int q = 0; q < RQ; ++q) {
            uint32_t acc = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w)
                acc = dot_traits<DT>::dot(x[w], qs[q * NW + w], acc);
            D[q * RG_TS + col] = rv ? dist_from<DT>(nx, qn[q], acc, g.mips) : __builtin_huge_valf();
        }
    }
};

// dynamic LDS: D[RQ][RG_TS] f32 | sd[RQ][k] f32 | sid[RQ][k] u32 | td[RQ] | tp[RQ] | tile extra
template <typename Tile>
__global__ __launch_bounds__(RG_THREADS) void compgt_refrule_kernel(RefGtArgs g, uint32_t is_u8)
{
    constexpr int RQ = Tile::RQ;
    extern __shared__ __align__(16) uint8_t lds[];
    float *D = reinterpret_cast<float *>(lds);
    float *sd = D + RQ * RG_TS;
    uint32_t *sid = reinterpret_cast<uint32_t *>(sd + (size_t)RQ * g.k);
    float *td = reinterpret_cast<float *>(sid + (size_t)RQ * g.k);
    uint32_t *tp = reinterpret_cast<uint32_t *>(td + RQ);
    void *extra = tp + RQ;
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint64_t q0 = (uint64_t)blockIdx.x * RQ;
    const uint32_t k = g.k;
    Tile tile;
    tile.init(g, q0, is_u8, extra, lane);
    if (threadIdx.x < RQ) {
        td[threadIdx.x] = -__FLT_MAX__;
        tp[threadIdx.x] = 0;
    }
    for (uint64_t c0 = 0; c0 < g.n; c0 += RG_CH) {
        __syncthreads(); // previous replay done with D
        tile.fill(g, c0, D, wv, lane);
        __syncthreads();
        for (uint32_t j = 0; j < RQ / 4; ++j) {
            const uint32_t q = wv * (RQ / 4) + j;
            if (q0 + q >= g.nq)
                break;
            float *qd = sd + (size_t)q * k;
            uint32_t *qi = sid + (size_t)q * k;
            float topd = td[q];
            uint32_t topp = tp[q];
            for (uint32_t gr = 0; gr < RG_CH / WAVE; ++gr) {
                const uint64_t r0 = c0 + gr * WAVE;
                if (r0 >= g.n)
                    break;
                const uint64_t row = r0 + lane;
                const float v = D[q * RG_TS + gr * WAVE + lane];
                if (r0 < k) { // fill the first k slots (CompGT.cpp:69-78)
                    if (row < k) {
                        qd[row] = v;
                        qi[row] = (uint32_t)row;
                    }
                    wave_sync();
                    if (r0 + WAVE >= k) {
                        const uint64_t key = first_max(qd, k, qd[0], lane);
                        topd = unord_f((uint32_t)(key >> 32));
                        topp = ~(uint32_t)key;
                    }
                }
                uint64_t m = ballot(row >= k && v < topd);
                while (m) {
                    const uint32_t p = (uint32_t)__builtin_ctzll(m);
                    const float d = __shfl(v, p);
                    const float nd = qd[0]; // read before the eviction (CompGT.cpp:82-84)
                    wave_sync();
                    if (lane == 0) {
                        qd[topp] = d;
                        qi[topp] = (uint32_t)(r0 + p);
                    }
                    wave_sync();
                    const uint64_t key = first_max(qd, k, nd, lane);
                    topd = unord_f((uint32_t)(key >> 32));
                    topp = ~(uint32_t)key;
                    m = ballot(row >= k && v < topd && lane > p);
                }
            }
            if (lane == 0) {
                td[q] = topd;
                tp[q] = topp;
            }
        }
    }
    __syncthreads();
    // sort each query's slots by (dist, id); D is free now (RQ/4 x 256 keys fit)
    uint32_t n2 = 1;
    while (n2 < k)
        n2 <<= 1;
    uint64_t *keys = reinterpret_cast<uint64_t *>(D) + (size_t)wv * 256;
    for (uint32_t j = 0; j < RQ / 4; ++j) {
        const uint32_t q = wv * (RQ / 4) + j;
        if (q0 + q >= g.nq)
            break;
        for (uint32_t i = lane; i < n2; i += WAVE)
            keys[i] = i < k ? make_key(sd[(size_t)q * k + i], sid[(size_t)q * k + i]) : ~0ull;
        wave_sync();
        rg_wave_bitonic(keys, n2, lane);
        for (uint32_t i = lane; i < k; i += WAVE) {
            g.ids[(q0 + q) * k + i] = key_id(keys[i]);
            if (g.dists)
                g.dists[(q0 + q) * k + i] = key_dist(keys[i]);
        }
        wave_sync();
    }
}

// exact top-k over base rows [r0, r1) of slice blockIdx.y: RQ queries per
// workgroup; per query a sorted key list top[q][0:cnt) in LDS
template <typename Tile>
__global__ __launch_bounds__(RG_THREADS) void compgt_tile_topk_kernel(RefGtArgs g, uint32_t is_u8,
                                                                      uint64_t slice_rows,
                                                                      uint32_t slices,
                                                                      uint64_t *partial)
{
    constexpr int RQ = Tile::RQ;
    extern __shared__ __align__(16) uint8_t lds[];
    float *D = reinterpret_cast<float *>(lds);
    uint64_t *top = reinterpret_cast<uint64_t *>(D + RQ * RG_TS);
    uint32_t *tn = reinterpret_cast<uint32_t *>(top + (size_t)RQ * g.k);
    void *extra = tn + RQ;
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint64_t q0 = (uint64_t)blockIdx.x * RQ;
    const uint32_t k = g.k;
    const uint64_t r0 = (uint64_t)blockIdx.y * slice_rows;
    const uint64_t r1 = min(g.n, r0 + slice_rows);
    Tile tile;
    tile.init(g, q0, is_u8, extra, lane);
    if (threadIdx.x < RQ)
        tn[threadIdx.x] = 0;
    for (uint64_t c0 = r0; c0 < r1; c0 += RG_CH) {
        __syncthreads(); // the previous chunk's lists are done with D
        tile.fill(g, c0, D, wv, lane);
        __syncthreads();
        for (uint32_t j = 0; j < RQ / 4; ++j) {
            const uint32_t q = wv * (RQ / 4) + j;
            if (q0 + q >= g.nq)
                break;
            uint64_t *tq = top + (size_t)q * k;
            uint32_t cnt = tn[q];
            uint64_t thr = cnt < k ? ~0ull : tq[k - 1];
            float thr_d = cnt < k ? __builtin_huge_valf() : key_dist(thr);
            for (uint32_t gr = 0; gr < RG_CH / WAVE; ++gr) {
                const uint64_t row = c0 + gr * WAVE + lane;
                const bool rv = row < r1;
                const float v = D[q * RG_TS + gr * WAVE + lane];
                // a superset by distance, then the exact (dist, id) test
                uint64_t m = ballot(rv && v <= thr_d);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    const uint64_t key = make_key(__shfl(v, l), (uint32_t)(c0 + gr * WAVE + l));
                    if (key >= thr)
                        continue;
                    // insert at pos = #entries below key; the list grows to k
                    uint32_t c = 0;
                    for (uint32_t i = lane; i < cnt; i += WAVE)
                        c += tq[i] < key;
                    const uint32_t pos = wave_sum(c);
                    const uint32_t ncnt = min(cnt + 1, k);
                    for (int32_t t = (int32_t)ncnt - 2; t >= (int32_t)pos; t -= WAVE) {
                        const int32_t idx = t - (int32_t)lane;
                        uint64_t x = 0;
                        if (idx >= (int32_t)pos)
                            x = tq[idx];
                        wave_sync();
                        if (idx >= (int32_t)pos)
                            tq[idx + 1] = x;
                        wave_sync();
                    }
                    if (lane == 0)
                        tq[pos] = key;
                    wave_sync();
                    cnt = ncnt;
                    if (cnt == k) {
                        thr = tq[k - 1];
                        thr_d = key_dist(thr);
                    }
                }
            }
            if (lane == 0)
                tn[q] = cnt;
            wave_sync();
        }
    }
    __syncthreads();
    for (uint32_t q = wv; q < RQ; q += RG_THREADS / WAVE) {
        if (q0 + q >= g.nq)
            break;
        const uint32_t cnt = tn[q];
        for (uint32_t i = lane; i < k; i += WAVE)
            partial[((q0 + q) * slices + blockIdx.y) * k + i] =
                i < cnt ? top[(size_t)q * k + i] : ~0ull;
    }
}

template <int RQ> static size_t tile_topk_lds(uint32_t k, size_t extra)
{
    return (size_t)RQ * RG_TS * 4 + (size_t)RQ * k * 8 + RQ * 4 + extra;
}

template <typename Tile>
static int launch_tile_topk(const RefGtArgs &g, uint32_t is_u8, size_t extra, hipStream_t s)
{
    constexpr int RQ = Tile::RQ;
    const size_t lds = tile_topk_lds<RQ>(g.k, extra);
    if (lds > (size_t)g_dev.lds_max)
        return set_error(NBPG_EUNSUP, "tile ground truth: k = %u too large", g.k);
    // slices of whole 256-row chunks: enough workgroups to fill the device,
    // at most 8192 merged keys per query
    const uint64_t groups = (g.nq + RQ - 1) / RQ;
    uint64_t slices = std::max<uint64_t>(1, ((uint64_t)g_dev.cu * 4 + groups - 1) / groups);
    slices = std::min<uint64_t>(slices, std::max<uint64_t>(1, 8192 / g.k));
    slices = std::min<uint64_t>(slices, (g.n + RG_CH - 1) / RG_CH);
    const uint64_t chunks = (g.n + RG_CH - 1) / RG_CH;
    const uint64_t slice_rows = (chunks + slices - 1) / slices * RG_CH;
    slices = (g.n + slice_rows - 1) / slice_rows;
    uint64_t *partial = nullptr;
    if (hipMallocAsync((void **)&partial, g.nq * slices * g.k * 8, s) != hipSuccess)
        return set_error(NBPG_ENOMEM, "tile ground truth partial lists");
    hipLaunchKernelGGL((compgt_tile_topk_kernel<Tile>), dim3((uint32_t)groups, (uint32_t)slices),
                       dim3(RG_THREADS), lds, s, g, is_u8, slice_rows, (uint32_t)slices, partial);
    int rc = hipGetLastError() == hipSuccess ? NBPG_OK
                                             : set_error(NBPG_EHIP, "tile ground truth launch");
    if (rc == NBPG_OK)
        rc = compgt_merge_launch(partial, g.nq, (uint32_t)slices, g.k, g.ids, g.dists, s);
    hipFreeAsync(partial, s);
    return rc;
}

// the exact (dist, id) top-k from distance tiles (MFMA tiles when use_mfma)
int compgt_tiles(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                 const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k, bool use_mfma,
                 uint32_t mips, uint32_t *ids, float *dists, hipStream_t s)
{
    RefGtArgs g;
    g.row_stride = row_stride;
    g.base = base;
    g.queries = queries;
    g.n = n;
    g.nq = nq;
    g.dims = dims;
    g.k = k;
    g.ids = ids;
    g.dists = dists;
    g.mips = mips;
    if (use_mfma) {
        // the most queries per workgroup whose tile and lists fit the LDS:
        // every query group streams the whole base once
        const uint32_t is_u8 = dtype == NBPG_U8;
        // ... while at least three workgroups still share a CU's LDS
        const size_t budget = (size_t)g_dev.lds_max / 3;
        const int qt = tile_topk_lds<64>(k, 0) <= budget ? 4 : (tile_topk_lds<32>(k, 0) <= budget ? 2 : 1);
        auto go = [&](auto KS) -> int {
            constexpr int ks = decltype(KS)::value;
            if (qt == 4)
                return launch_tile_topk<MfmaTile<ks, 4>>(g, is_u8, 0, s);
            if (qt == 2)
                return launch_tile_topk<MfmaTile<ks, 2>>(g, is_u8, 0, s);
            return launch_tile_topk<MfmaTile<ks, 1>>(g, is_u8, 0, s);
        };
        switch ((dims + 63) / 64) {
        case 1: return go(std::integral_constant<int, 1>());
        case 2: return go(std::integral_constant<int, 2>());
        case 3: return go(std::integral_constant<int, 3>());
        default: return go(std::integral_constant<int, 4>());
        }
    }
    const uint32_t nch = nch_for_bytes(dims);
    int rc = NBPG_OK;
    int drc = dispatch_dt_nch(dtype, nch, [&](auto DT, auto NCH) {
        using T = DotTile<DT.value, NCH.value, 1>;
        rc = launch_tile_topk<T>(g, 0, (size_t)T::RQ * (T::NW + 1) * 4, s);
    });
    return drc != NBPG_OK ? drc : rc;
}

template <typename Tile>
static int launch_refrule(const RefGtArgs &g, uint32_t is_u8, size_t extra, hipStream_t s)
{
    constexpr int RQ = Tile::RQ;
    const size_t lds = (size_t)RQ * RG_TS * 4 + (size_t)RQ * g.k * 8 + RQ * 8 + extra;
    if (lds > 65536 || (size_t)RQ * RG_TS * 4 < 4 * 256 * 8)
        return set_error(NBPG_EUNSUP, "reference-rule ground truth: k = %u too large", g.k);
    const uint64_t groups = (g.nq + RQ - 1) / RQ;
    hipLaunchKernelGGL((compgt_refrule_kernel<Tile>), dim3((uint32_t)groups), dim3(RG_THREADS),
                       lds, s, g, is_u8);
    if (hipGetLastError() != hipSuccess)
        return set_error(NBPG_EHIP, "reference-rule ground truth launch");
    return NBPG_OK;
}

// use_mfma: u8 / i8 rows of a multiple of 16 bytes (<= 256), 16-byte aligned
int compgt_refrule(int dtype, const uint8_t *base, uint64_t n, const uint8_t *queries, uint64_t nq,
                   uint32_t dims, uint32_t k, bool use_mfma, uint32_t mips, uint32_t *ids,
                   float *dists, hipStream_t s)
{
    RefGtArgs g;
    g.mips = mips;
    g.base = base;
    g.queries = queries;
    g.n = n;
    g.nq = nq;
    g.dims = dims;
    g.k = k;
    g.ids = ids;
    g.dists = dists;
    g.row_stride = dims;
    const bool small = k <= 96; // 32 queries per workgroup still fit 64 KB of LDS
    if (use_mfma) {
        const uint32_t is_u8 = dtype == NBPG_U8;
        switch ((dims + 63) / 64) {
        case 1: return small ? launch_refrule<MfmaTile<1, 2>>(g, is_u8, 0, s) : launch_refrule<MfmaTile<1, 1>>(g, is_u8, 0, s);
        case 2: return small ? launch_refrule<MfmaTile<2, 2>>(g, is_u8, 0, s) : launch_refrule<MfmaTile<2, 1>>(g, is_u8, 0, s);
        case 3: return small ? launch_refrule<MfmaTile<3, 2>>(g, is_u8, 0, s) : launch_refrule<MfmaTile<3, 1>>(g, is_u8, 0, s);
        default: return small ? launch_refrule<MfmaTile<4, 2>>(g, is_u8, 0, s) : launch_refrule<MfmaTile<4, 1>>(g, is_u8, 0, s);
        }
    }
    const uint32_t nch = nch_for_bytes(dims);
    int rc = NBPG_OK;
    int drc = dispatch_dt_nch(dtype, nch, [&](auto DT, auto NCH) {
        // 16 queries per workgroup: the query words share the 64 KB with the slots
        using T = DotTile<DT.value, NCH.value, 1>;
        rc = launch_refrule<T>(g, 0, (size_t)T::RQ * (T::NW + 1) * 4, s);
    });
    return drc != NBPG_OK ? drc : rc;
}

} // namespace nbpg
