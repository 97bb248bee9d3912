// compgt_mfma.hip -- exact brute-force top-k on int8 matrix cores.
//
// Same result as compgt_plain (the exact (dist, id) top-k of every query,
// what compute_groundtruth at CompGT.cpp:29-134 aims at), three passes:
//   1. threshold: the plain scan over S base rows sampled at a fixed stride
//      gives each query's k-th smallest distance there.  The k-th smallest
//      distance of a subset is never below the global k-th smallest, so every
//      true top-k row has a distance <= this threshold.
//   2. scan: v_mfma_i32_16x16x64_i8 Gram tiles.  u8 rows enter as x - 128
//      (x ^ 0x80 per byte); |x-q|^2 = |x'-q'|^2 = |x'|^2 + |q'|^2 - 2<x',q'>
//      is exact in int32.  A wave keeps 16*QW queries' fragments in
//      registers and streams 16-row base blocks; every row within its
//      query's threshold is appended to that query's candidate list.
//   3. select: per query, a bitonic sort of its candidates; a query whose
//      list overflowed is recomputed by the plain kernel.
#include <algorithm>
#include <climits>
#include <vector>

#include "internal.h"

namespace nbpg {

typedef int v4i __attribute__((ext_vector_type(4)));

#define MF_CAND 4096 // candidate slots per query
#define MF_QW 2      // 16-query A tiles per wave (2: 106 VGPRs, 4 waves per SIMD)
#define MF_WAVES 4

int compgt_plain_launch(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                        const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k,
                        uint32_t mips, uint32_t *ids, float *dists, hipStream_t s);
int compgt_tiles(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                 const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k, bool use_mfma,
                 uint32_t mips, uint32_t *ids, float *dists, hipStream_t s);

// per row, with x' = x ^ 0x80 = x - 128 (u8) or x (i8), over dims bytes:
// L2: |x'|^2.  Inner product: half of what turns <x',q'> back into <x,q>,
// <x,q> = <x',q'> + A(x) + A(q) with A(x) = 128 sum(x') + 8192 dims (u8), 0 (i8)
// rows of dims bytes (a multiple of 16, 16-byte aligned): 8 lanes per row,
// each summing 16-byte pieces 128 bytes apart, then an xor reduction over
// the 8 (integer sums: the split is exact)
__global__ void mf_norms_kernel(const uint8_t *__restrict__ rows, uint64_t n, uint32_t dims,
                                int is_u8, uint32_t mips, uint32_t *__restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t i = t / 8;
    const uint32_t piece = (uint32_t)(t % 8);
    const uint32_t xm = is_u8 ? 0x80808080u : 0u;
    int acc = 0;
    if (i < n)
        for (uint32_t off = 16 * piece; off < dims; off += 128) {
            const uint4 u = *reinterpret_cast<const uint4 *>(rows + i * dims + off);
            const uint32_t w[4] = {u.x ^ xm, u.y ^ xm, u.z ^ xm, u.w ^ xm};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc = __builtin_amdgcn_sdot4((int)w[j], mips ? 0x01010101 : (int)w[j], acc, false);
        }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    acc += __shfl_xor(acc, 4);
    if (i < n && piece == 0) {
        if (mips)
            acc = is_u8 ? 128 * acc + 8192 * (int)dims : 0;
        out[i] = (uint32_t)acc;
    }
}

// threshold key: the sample's k-th distance with the largest id, so every
// row at that distance or closer passes `key <= thr`
__global__ void mf_thresholds_kernel(const float *__restrict__ d, uint64_t nq, uint32_t k,
                                     uint64_t *__restrict__ thr, uint32_t *__restrict__ cnt,
                                     uint32_t *__restrict__ ovf)
{
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    thr[q] = make_key(d[q * k + k - 1], 0xFFFFFFFFu);
    cnt[q] = 0;
    ovf[q] = 0;
}

// rows[i] = src[idx[i]] (dims bytes each)
__global__ void mf_gather_kernel(const uint8_t *__restrict__ src, const uint32_t *__restrict__ idx,
                                 uint32_t m, uint32_t dims, uint8_t *__restrict__ rows)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)m * dims)
        return;
    const uint64_t i = t / dims, b = t % dims;
    rows[t] = src[(uint64_t)idx[i] * dims + b];
}

// out[idx[i]] = compact[i] (k results each)
__global__ void mf_scatter_kernel(const uint32_t *__restrict__ cids, const float *__restrict__ cd,
                                  const uint32_t *__restrict__ idx, uint32_t m, uint32_t k,
                                  uint32_t *__restrict__ ids, float *__restrict__ dists)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)m * k)
        return;
    const uint64_t i = t / k, j = t % k, o = (uint64_t)idx[i] * k + j;
    ids[o] = cids[t];
    if (dists)
        dists[o] = cd[t];
}

// 64-row blocks of the slice pass through LDS: the workgroup's 256 threads
// load a block once (KS 16-byte pieces each, into registers while the
// previous block is scored, then into the other LDS buffer) and every
// wavefront reads its MFMA operands from there, so a base row is read from
// memory once per 16 * MF_QW * MF_WAVES queries.
#define MF_BLK 64

template <int KS>
__global__ __launch_bounds__(64 * MF_WAVES) void mf_scan_kernel(
    const uint8_t *__restrict__ base, uint64_t n, const uint8_t *__restrict__ queries, uint64_t nq,
    uint32_t dims, int is_u8, const uint32_t *__restrict__ nx, const uint32_t *__restrict__ nqv,
    const uint64_t *__restrict__ thr, uint64_t slice_rows, uint32_t *__restrict__ cnt,
    uint64_t *__restrict__ cand, uint32_t mips)
{
    constexpr uint32_t PITCH = 64 * KS + 16; // row pitch in LDS (16-byte pad against bank conflicts)
    __shared__ __align__(16) uint8_t rows_lds[2][MF_BLK * PITCH];
    __shared__ int32_t xs_lds[2][MF_BLK];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint32_t row16 = lane % 16, kq = lane / 16;
    const uint64_t q0 = ((uint64_t)blockIdx.x * MF_WAVES + wv) * (16 * MF_QW);
    const bool active = q0 < nq; // inactive waves still load and synchronise
    const uint32_t xmask = is_u8 ? 0x80808080u : 0u;
    // query fragments: lane supplies 16 bytes of query row (l % 16) at k-chunk (l / 16)
    v4i a[MF_QW][KS];
    // the admission test in integers: with T the threshold distance (an
    // exact integer), L2 admits xn - 2c <= T - qn, the inner product
    // c + xn >= -T - qn; U holds the right-hand sides (never true for a
    // missing query)
    int32_t U[MF_QW][4];
    uint32_t qn[MF_QW][4];
#pragma unroll
    for (int t = 0; t < MF_QW; ++t) {
        const uint64_t q = q0 + 16 * t + row16;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint32_t off = 64 * s + 16 * kq;
            v4i v = {0, 0, 0, 0};
            if (q < nq && off < dims) {
                const uint4 u = *reinterpret_cast<const uint4 *>(queries + q * dims + off);
                v = v4i{(int)(u.x ^ xmask), (int)(u.y ^ xmask), (int)(u.z ^ xmask), (int)(u.w ^ xmask)};
            }
            a[t][s] = v;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint64_t qr = q0 + 16 * t + 4 * kq + r;
            qn[t][r] = qr < nq ? nqv[qr] : 0u;
            const int32_t T = qr < nq ? (int32_t)key_dist(thr[qr]) : 0;
            U[t][r] = qr >= nq ? (mips ? INT_MAX : INT_MIN)
                               : (mips ? -T - (int32_t)qn[t][r] : T - (int32_t)qn[t][r]);
        }
    }
    const uint64_t r0 = (uint64_t)blockIdx.y * slice_rows;
    const uint64_t r1 = min(n, r0 + slice_rows);
    // this thread's pieces of a block: piece p = tid + 256 j is row p / (4 KS),
    // 16-byte chunk p % (4 KS)
    v4i stage[KS];
    int32_t xstage = 0;
    auto fetch = [&](uint64_t b0) {
#pragma unroll
        for (int j = 0; j < KS; ++j) {
            const uint32_t p = threadIdx.x + 64 * MF_WAVES * j;
            const uint64_t row = b0 + p / (4 * KS);
            const uint32_t off = 16 * (p % (4 * KS));
            v4i v = {0, 0, 0, 0};
            if (row < r1 && off < dims) {
                const uint4 u = *reinterpret_cast<const uint4 *>(base + row * dims + off);
                v = v4i{(int)(u.x ^ xmask), (int)(u.y ^ xmask), (int)(u.z ^ xmask), (int)(u.w ^ xmask)};
            }
            stage[j] = v;
        }
        if (threadIdx.x < MF_BLK) {
            const uint64_t row = b0 + threadIdx.x;
            // a missing row fails every test
            xstage = row < r1 ? (int32_t)nx[row] : (mips ? -(1 << 30) : (1 << 30));
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int j = 0; j < KS; ++j) {
            const uint32_t p = threadIdx.x + 64 * MF_WAVES * j;
            *reinterpret_cast<v4i *>(&rows_lds[buf][(p / (4 * KS)) * PITCH + 16 * (p % (4 * KS))]) =
                stage[j];
        }
        if (threadIdx.x < MF_BLK)
            xs_lds[buf][threadIdx.x] = xstage;
    };
    fetch(r0);
    put(0);
    __syncthreads();
    int buf = 0;
    for (uint64_t b0 = r0; b0 < r1; b0 += MF_BLK) {
        const bool more = b0 + MF_BLK < r1;
        if (more)
            fetch(b0 + MF_BLK); // in flight while this block is scored
        if (active) {
#pragma unroll
            for (int blk = 0; blk < MF_BLK / 16; ++blk) {
                const uint32_t rl = 16 * blk + row16;
                const uint64_t row = b0 + rl;
                v4i b[KS];
#pragma unroll
                for (int s = 0; s < KS; ++s)
                    b[s] = *reinterpret_cast<const v4i *>(&rows_lds[buf][rl * PITCH + 64 * s + 16 * kq]);
                const int32_t xs = xs_lds[buf][rl];
#pragma unroll
                for (int t = 0; t < MF_QW; ++t) {
                    v4i c = {0, 0, 0, 0};
#pragma unroll
                    for (int s = 0; s < KS; ++s)
                        c = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[t][s], b[s], c, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const bool adm = mips ? c[r] + xs >= U[t][r] : xs - 2 * c[r] <= U[t][r];
                        if (adm) {
                            const uint32_t xn = (uint32_t)xs;
                            const float d = mips ? -(float)(int32_t)((uint32_t)c[r] + qn[t][r] + xn)
                                                 : (float)(int32_t)(qn[t][r] + xn - 2u * (uint32_t)c[r]);
                            const uint64_t qr = q0 + 16 * t + 4 * kq + r;
                            const uint32_t p = atomicAdd(&cnt[qr], 1u);
                            if (p < MF_CAND)
                                cand[qr * MF_CAND + p] = make_key(d, (uint32_t)row);
                        }
                    }
                }
            }
        }
        if (more)
            put(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
}

// per query: sort the candidates, write k; flag overflowed queries
__global__ __launch_bounds__(256) void mf_select_kernel(const uint64_t *__restrict__ cand,
                                                       const uint32_t *__restrict__ cnt, uint32_t k,
                                                       uint32_t *__restrict__ ids,
                                                       float *__restrict__ dists,
                                                       uint32_t *__restrict__ overflow)
{
    __shared__ uint64_t s[MF_CAND];
    const uint64_t q = blockIdx.x;
    const uint32_t m = cnt[q];
    if (m > MF_CAND) {
        if (threadIdx.x == 0)
            overflow[q] = 1;
        return;
    }
    uint32_t n2 = 1;
    while (n2 < m)
        n2 <<= 1;
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x)
        s[i] = i < m ? cand[q * MF_CAND + i] : ~0ull;
    __syncthreads();
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = s[i], y = s[p];
                    if ((x > y) == ((i & kk) == 0)) {
                        s[i] = y;
                        s[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
        ids[q * k + i] = key_id(s[i]);
        if (dists)
            dists[q * k + i] = key_dist(s[i]);
    }
}

int compgt_mfma(int dtype, const uint8_t *base, uint64_t n, const uint8_t *queries, uint64_t nq,
                uint32_t dims, uint32_t k, uint32_t mips, uint32_t *out_ids, float *out_dists,
                hipStream_t s)
{
    if (dtype != NBPG_U8 && dtype != NBPG_I8)
        return set_error(NBPG_EUNSUP, "MFMA ground truth takes u8 or i8 rows");
    if (dims % 16 || dims > 256 || (reinterpret_cast<uintptr_t>(base) & 15) ||
        (reinterpret_cast<uintptr_t>(queries) & 15))
        return set_error(NBPG_EUNSUP, "MFMA ground truth needs 16-byte rows (dims %u)", dims);
    const int is_u8 = dtype == NBPG_U8;
    const uint32_t KS = (dims + 63) / 64;
    // 1. thresholds from every step-th row: about k * n / S rows then pass
    //    per query, kept to a quarter of the candidate slots
    const uint64_t want = std::max<uint64_t>(
        std::max<uint64_t>(131072, 64ull * k), (4ull * k * n + MF_CAND - 1) / MF_CAND);
    const uint64_t step = std::max<uint64_t>(1, n / want);
    const uint64_t S = (n + step - 1) / step; // >= min(n, want) >= k
    uint32_t *tid = nullptr, *cnt = nullptr, *nx = nullptr, *nqv = nullptr, *ovf = nullptr;
    float *td = nullptr;
    uint64_t *thr = nullptr, *cand = nullptr;
    auto cleanup = [&]() {
        hipFreeAsync(tid, s);
        hipFreeAsync(td, s);
        hipFreeAsync(cnt, s);
        hipFreeAsync(nx, s);
        hipFreeAsync(nqv, s);
        hipFreeAsync(ovf, s);
        hipFreeAsync(thr, s);
        hipFreeAsync(cand, s);
    };
    bool ok = hipMallocAsync((void **)&tid, nq * k * 4, s) == hipSuccess &&
              hipMallocAsync((void **)&td, nq * k * 4, s) == hipSuccess &&
              hipMallocAsync((void **)&cnt, nq * 4, s) == hipSuccess &&
              hipMallocAsync((void **)&nx, n * 4, s) == hipSuccess &&
              hipMallocAsync((void **)&nqv, nq * 4, s) == hipSuccess &&
              hipMallocAsync((void **)&ovf, nq * 4, s) == hipSuccess &&
              hipMallocAsync((void **)&thr, nq * 8, s) == hipSuccess &&
              hipMallocAsync((void **)&cand, nq * MF_CAND * 8, s) == hipSuccess;
    if (!ok) {
        cleanup();
        return set_error(NBPG_ENOMEM, "MFMA ground-truth workspace");
    }
    int rc = compgt_tiles(dtype, base, S, step * dims, queries, nq, dims, k, true, mips, tid, td, s);
    if (rc != NBPG_OK) {
        cleanup();
        return rc;
    }
    hipLaunchKernelGGL(mf_thresholds_kernel, dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, s,
                       td, nq, k, thr, cnt, ovf);
    hipLaunchKernelGGL(mf_norms_kernel, dim3((uint32_t)((n + 31) / 32)), dim3(256), 0, s, base, n,
                       dims, is_u8, mips, nx);
    hipLaunchKernelGGL(mf_norms_kernel, dim3((uint32_t)((nq + 31) / 32)), dim3(256), 0, s,
                       queries, nq, dims, is_u8, mips, nqv);
    // 2. MFMA scan: query tiles x base slices
    const uint64_t qpb = 16ull * MF_QW * MF_WAVES;
    const uint64_t tiles = (nq + qpb - 1) / qpb;
    uint64_t slices = std::max<uint64_t>(1, (uint64_t)g_dev.cu * 8 / tiles);
    slices = std::min<uint64_t>(slices, std::max<uint64_t>(1, n / 4096));
    const uint64_t slice_rows = ((n + slices - 1) / slices + MF_BLK - 1) / MF_BLK * MF_BLK;
    slices = (n + slice_rows - 1) / slice_rows;
    dim3 grid((uint32_t)tiles, (uint32_t)slices), block(64 * MF_WAVES);
    switch (KS) {
    case 1: hipLaunchKernelGGL((mf_scan_kernel<1>), grid, block, 0, s, base, n, queries, nq, dims, is_u8, nx, nqv, thr, slice_rows, cnt, cand, mips); break;
    case 2: hipLaunchKernelGGL((mf_scan_kernel<2>), grid, block, 0, s, base, n, queries, nq, dims, is_u8, nx, nqv, thr, slice_rows, cnt, cand, mips); break;
    case 3: hipLaunchKernelGGL((mf_scan_kernel<3>), grid, block, 0, s, base, n, queries, nq, dims, is_u8, nx, nqv, thr, slice_rows, cnt, cand, mips); break;
    default: hipLaunchKernelGGL((mf_scan_kernel<4>), grid, block, 0, s, base, n, queries, nq, dims, is_u8, nx, nqv, thr, slice_rows, cnt, cand, mips); break;
    }
    if (hipGetLastError() != hipSuccess) {
        cleanup();
        return set_error(NBPG_EHIP, "MFMA scan launch");
    }
    // 3. select, then redo overflowed queries with the plain scan
    hipLaunchKernelGGL(mf_select_kernel, dim3((uint32_t)nq), dim3(256), 0, s, cand, cnt, k, out_ids,
                       out_dists, ovf);
    std::vector<uint32_t> h(nq);
    if (hipMemcpyAsync(h.data(), ovf, nq * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        return set_error(NBPG_EHIP, "overflow readback");
    }
    std::vector<uint32_t> redo;
    for (uint64_t q = 0; q < nq; ++q)
        if (h[q])
            redo.push_back((uint32_t)q);
    if (!redo.empty()) {
        const uint32_t m = (uint32_t)redo.size();
        uint32_t *idx = nullptr, *cids = nullptr;
        float *cd = nullptr;
        uint8_t *rows = nullptr;
        ok = hipMallocAsync((void **)&idx, m * 4ull, s) == hipSuccess &&
             hipMallocAsync((void **)&cids, (uint64_t)m * k * 4, s) == hipSuccess &&
             hipMallocAsync((void **)&cd, (uint64_t)m * k * 4, s) == hipSuccess &&
             hipMallocAsync((void **)&rows, (uint64_t)m * dims, s) == hipSuccess &&
             hipMemcpyAsync(idx, redo.data(), m * 4ull, hipMemcpyHostToDevice, s) == hipSuccess;
        if (ok) {
            const uint64_t tb = (uint64_t)m * dims, tk = (uint64_t)m * k;
            hipLaunchKernelGGL(mf_gather_kernel, dim3((uint32_t)((tb + 255) / 256)), dim3(256), 0,
                               s, queries, idx, m, dims, rows);
            rc = compgt_plain_launch(dtype, base, n, dims, rows, m, dims, k, mips, cids, cd, s);
            if (rc == NBPG_OK)
                hipLaunchKernelGGL(mf_scatter_kernel, dim3((uint32_t)((tk + 255) / 256)),
                                   dim3(256), 0, s, cids, cd, idx, m, k, out_ids, out_dists);
            // the host vector redo must outlive the copy
            if (hipStreamSynchronize(s) != hipSuccess && rc == NBPG_OK)
                rc = set_error(NBPG_EHIP, "overflow redo");
        } else {
            rc = set_error(NBPG_ENOMEM, "overflow redo workspace");
        }
        hipFreeAsync(idx, s);
        hipFreeAsync(cids, s);
        hipFreeAsync(cd, s);
        hipFreeAsync(rows, s);
        if (!ok)
            hipStreamSynchronize(s);
    }
    cleanup();
    return rc;
}

} // namespace nbpg
