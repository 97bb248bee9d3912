// compgt_mfma.hip -- exact brute-force top-k on int8 matrix cores.
//
// Same result as compgt_plain (the exact (dist, id) top-k of every query,
// what compute_groundtruth at CompGT.cpp:29-134 aims at), in four passes
// per chunk of queries, all on one stream with no host synchronisation and
// no allocation once the per-device workspace has grown to the call's size:
//   1. threshold: v_mfma_i32_16x16x64_i8 distance tiles over S base rows
//      sampled at a fixed stride.  Each lane keeps, per query, the MF_TV (2)
//      smallest distances of the rows it sees (a quarter of each 16-row
//      block); a small kernel takes the J-th smallest of all those values as
//      the query's bound U.  They are distances of distinct rows, so at
//      least J rows lie within U, and about J * stride rows of the base
//      (rarely more: a lane would have to see more than two of the query's
//      J nearest sample rows).
//   2. scan: the same tiles over every base row; a row is kept when its
//      distance is within U.  The row term enters the matrix core as the
//      accumulator's start value, so the common case per 16 x 16 tile is
//      one compare on the largest of a lane's four values and one
//      wave-wide branch.  Kept rows go to a segment of their own per
//      (query, base slice, lane), counted in a register: no atomics and no
//      memory round trip in the scan.  A kept row is one 64-bit store of
//      (partial distance as an order-preserving unsigned, row id).
//   3. select: per query, its segments gathered, cut to the rows up to the
//      histogram bucket of the k-th distance, and bitonic-sorted.  They
//      hold every row within U, so when they hold at least k rows (and no
//      segment overflowed) the first k are the exact top-k, ties at the
//      k-th distance included.  Otherwise the query goes on a device list ...
//   4. ... that the plain scan redoes (a launch sized for the whole chunk
//      that reads the list's length on the device; empty almost always).
// Rows enter as x - 128 (x ^ 0x80 per byte, u8) or x (i8):
// |x-q|^2 = |x'|^2 + |q'|^2 - 2<x',q'>, exact in int32.  Inner product:
// <x,q> = <x',q'> + A(x) + A(q) (mf_norms_kernel).
#include <algorithm>
#include <climits>
#include <mutex>
#include <type_traits>
#include <vector>

#include "internal.h"

namespace nbpg {

typedef int v4i __attribute__((ext_vector_type(4)));

#ifndef MF_QW
#define MF_QW 4 // 16-query tiles per wavefront (B operand, registers)
#endif
// wavefronts per workgroup: 16 (1024 queries share each staged block) for
// rows of up to 128 bytes, whose scan fits 128 VGPRs (4 waves per SIMD, one
// workgroup per CU); 8 for wider rows, whose scan holds 2 waves per SIMD
__host__ __device__ constexpr uint32_t mf_waves(uint32_t KS) { return KS <= 2 ? 16u : 8u; }
#ifndef MF_BLK
#define MF_BLK 128 // base rows per block staged through LDS
#endif
#define MF_J 16 // bound = the J-th smallest of the per-lane top-MF_TV sample values
#define MF_TV 2 // smallest sample distances each lane keeps per query
#define MF_QCHUNK 16384
#define MF_MAX_TSLICES 64

uint64_t compgt_plain_partial_bytes(uint64_t nq, uint64_t n, uint32_t k, bool listed);
int compgt_plain_core(int dtype, const uint8_t *base, uint64_t n, uint64_t row_stride,
                      const uint8_t *queries, uint64_t nq, uint32_t dims, uint32_t k,
                      uint32_t mips, uint32_t *ids, float *dists, hipStream_t s,
                      uint64_t *partial, const uint32_t *qidx, const uint32_t *qcount);

// per row, with x' = x ^ 0x80 = x - 128 (u8) or x (i8), over dims bytes:
// L2: |x'|^2.  Inner product: A(x) = 128 sum(x') + 8192 dims (u8), 0 (i8),
// the part of <x,q> = <x',q'> + A(x) + A(q) that belongs to the row.
// Rows of dims bytes (a multiple of 16, 16-byte aligned): 8 lanes per row,
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

struct MfArgs {
    const uint8_t *base;
    uint64_t rows;       // rows of this pass (sample rows, or the whole base)
    uint64_t row_step;   // pass row i is base row i * row_step
    const uint8_t *queries;
    uint64_t nq;
    uint32_t dims;
    int is_u8;
    const uint32_t *nx;  // per base row: |x'|^2, or A(x)
    const uint32_t *nqv; // per query: |q'|^2, or A(q)
    uint64_t slice_rows;
    const int32_t *U;    // scan: per-query bound on the partial distance
    uint64_t *cand;      // scan: segments of seg_cap keys per (query, slice, lane group)
    uint32_t *segcnt;    // scan: rows kept per segment (beyond seg_cap: overflowed)
    uint32_t seg_cap;
    int32_t *tv;         // threshold pass: per query and slice, 4 lanes x MF_TV values
};

// Distances are carried as the partial distance e = d - |q'|^2 (L2:
// |x'|^2 - 2c) or e = d + A(q) (inner product: -c - A(x)), an exact int32;
// per query the order of e is the order of d.  Each block row's w holds
// |x'|^2 or -A(x); a missing row has w = 2^30, beyond every bound.
template <int KS, bool MIPS, bool THRESH>
__global__ __launch_bounds__(64 * mf_waves(KS)) __attribute__((amdgpu_waves_per_eu(KS <= 2 ? 4 : 2)))
void mf_scan_kernel(MfArgs a)
{
    constexpr uint32_t PITCH = 64 * KS + 16; // LDS row pitch (16-byte pad against bank conflicts)
    constexpr int32_t MISSING = 1 << 30;
    __shared__ __align__(16) uint8_t rows_lds[2][MF_BLK * PITCH];
    __shared__ __align__(16) int32_t w_lds[2][MF_BLK];
    __shared__ __align__(16) int32_t i_lds[2][MF_BLK];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint32_t col = lane % 16, kq = lane / 16;
    constexpr uint32_t WAVES = mf_waves(KS);
    const uint64_t q0 = ((uint64_t)blockIdx.x * WAVES + wv) * (16 * MF_QW);
    const bool active = q0 < a.nq; // inactive waves still stage blocks and synchronise
    const uint32_t xmask = a.is_u8 ? 0x80808080u : 0u;
    // D = rows x queries: the A operand is a 16-row block (lane: row `col`,
    // 16 bytes at k-chunk kq), the B operand 16 queries (lane: query `col`,
    // the same chunk), and lane l holds D[4 kq + r][col], r = 0..3
    v4i qf[MF_QW][KS];
    int32_t V[MF_QW], mt[MF_QW][MF_TV]; // threshold: ascending
    uint32_t kept[MF_QW]; // scan: rows this lane kept per query tile
    uint32_t segb[MF_QW]; // scan: first key of this lane's segment per query tile
#pragma unroll
    for (int t = 0; t < MF_QW; ++t) {
        const uint64_t q = q0 + 16 * t + col;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint32_t off = 64 * s + 16 * kq;
            v4i v = {0, 0, 0, 0};
            if (q < a.nq && off < a.dims) {
                const uint4 u = *reinterpret_cast<const uint4 *>(a.queries + q * a.dims + off);
                v = v4i{(int)(u.x ^ xmask), (int)(u.y ^ xmask), (int)(u.z ^ xmask),
                        (int)(u.w ^ xmask)};
            }
            qf[t][s] = v;
        }
        // the scan's test on c' (below): c' >= V, V = -floor(U / 2) (L2) or
        // -U (inner product); never true for a missing query
        const int32_t u = q < a.nq ? a.U[q] : 0;
        V[t] = THRESH || q >= a.nq ? INT_MAX : (MIPS ? -u : -(u >> 1));
#pragma unroll
        for (int i = 0; i < MF_TV; ++i)
            mt[t][i] = INT_MAX;
        kept[t] = 0;
        // the host keeps every segment key index below 2^32
        segb[t] = (uint32_t)(((q * gridDim.y + blockIdx.y) * 4 + kq) * a.seg_cap);
    }
    const uint64_t r0 = (uint64_t)blockIdx.y * a.slice_rows;
    const uint64_t r1 = min(a.rows, r0 + a.slice_rows);
    // this thread's pieces of a block: piece p = tid + 64 WAVES j is row
    // p / (4 KS), 16-byte chunk p % (4 KS).  Two register sets: a block is
    // loaded two blocks ahead of its use, so its load has two blocks' worth
    // of scoring to arrive before it is written to LDS.
    constexpr uint32_t PIECES = (MF_BLK * 4 * KS + 64 * WAVES - 1) / (64 * WAVES);
    v4i stage[2][PIECES];
    int32_t wstage[2] = {MISSING, MISSING};
    const uint64_t row_bytes = a.row_step * a.dims;
    const uint32_t row_bytes32 = (uint32_t)row_bytes;
    // register set R (a compile-time index: register arrays take no
    // run-time index) <- block b0; the block's base address is uniform, each
    // piece adds its row and chunk offset
    auto fetch = [&](uint64_t b0, auto R) {
        const uint8_t *bb = a.base + b0 * row_bytes;
        const uint64_t left = b0 < r1 ? r1 - b0 : 0; // rows of this block that exist
#pragma unroll
        for (uint32_t j = 0; j < PIECES; ++j) {
            const uint32_t p = threadIdx.x + 64 * WAVES * j;
            const uint32_t prow = p / (4 * KS), off = 16 * (p % (4 * KS));
            v4i v = {0, 0, 0, 0};
            if (p < MF_BLK * 4 * KS && off < a.dims && prow < left) {
                // a block spans under 2^31 bytes (the host bounds the
                // sample stride), so the offset fits 32 bits
                const uint4 u = *reinterpret_cast<const uint4 *>(bb + (prow * row_bytes32 + off));
                v = v4i{(int)(u.x ^ xmask), (int)(u.y ^ xmask), (int)(u.z ^ xmask),
                        (int)(u.w ^ xmask)};
            }
            stage[R.value][j] = v;
        }
        if (threadIdx.x < MF_BLK) {
            const bool ok = threadIdx.x < left;
            const int32_t x = ok ? (int32_t)a.nx[b0 * a.row_step + threadIdx.x * a.row_step] : 0;
            wstage[R.value] = ok ? (MIPS ? -x : x) : MISSING;
        }
    };
    // LDS buffer buf <- register set R; each row's w and the accumulator's
    // start value -floor(w / 2) (L2) or -w (inner product)
    auto put = [&](int buf, auto R) {
#pragma unroll
        for (uint32_t j = 0; j < PIECES; ++j) {
            const uint32_t p = threadIdx.x + 64 * WAVES * j;
            if (p < MF_BLK * 4 * KS)
                *reinterpret_cast<v4i *>(&rows_lds[buf][(p / (4 * KS)) * PITCH + 16 * (p % (4 * KS))]) =
                    stage[R.value][j];
        }
        if (threadIdx.x < MF_BLK) {
            const int32_t w = wstage[R.value];
            w_lds[buf][threadIdx.x] = w;
            i_lds[buf][threadIdx.x] = THRESH ? 0 : (MIPS ? -w : -(w >> 1));
        }
    };
    // every query tile against the block in LDS buffer buf (rows b0..)
    auto score = [&](int buf, uint64_t b0) {
        if (!active)
            return;
        v4i b[KS], init;
        auto ldb = [&](int blk) {
#pragma unroll
            for (int s = 0; s < KS; ++s)
                b[s] = *reinterpret_cast<const v4i *>(
                    &rows_lds[buf][(16 * blk + col) * PITCH + 64 * s + 16 * kq]);
            init = *reinterpret_cast<const v4i *>(&i_lds[buf][16 * blk + 4 * kq]);
        };
        ldb(0);
#pragma unroll
        for (int blk = 0; blk < MF_BLK / 16; ++blk) {
            // scan: the accumulator starts at -h (h = floor(w / 2), L2) or
            // -w (inner product), shared by every query tile, so the matrix
            // core returns c' = c - h or c - w.  e = w - 2c <= U implies
            // c' >= -floor(U / 2) (a superset that adds only e = U + 1), and
            // e = w - c <= U is c' >= -U: one compare on the largest of the
            // four values per tile
            // chunk-major order: each accumulator's next product is issued
            // MF_QW products after its last, so none waits on its own
            // result; one wave-wide branch per sub-block
            v4i c[MF_QW];
#pragma unroll
            for (int t = 0; t < MF_QW; ++t)
                c[t] = init;
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int t = 0; t < MF_QW; ++t)
                    c[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(b[s], qf[t][s], c[t], 0, 0, 0);
            // the next sub-block's fragments, in flight during this one's test
            if (blk + 1 < MF_BLK / 16)
                ldb(blk + 1);
            if (THRESH) {
                const v4i w4 = *reinterpret_cast<const v4i *>(&w_lds[buf][16 * blk + 4 * kq]);
#pragma unroll
                for (int t = 0; t < MF_QW; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        // insert e into the ascending MF_TV smallest
                        int32_t e = MIPS ? w4[r] - c[t][r] : w4[r] - 2 * c[t][r];
#pragma unroll
                        for (int i = 0; i < MF_TV; ++i) {
                            const int32_t lo = min(mt[t][i], e);
                            e = max(mt[t][i], e);
                            mt[t][i] = lo;
                        }
                    }
            } else {
                bool hit = false;
#pragma unroll
                for (int t = 0; t < MF_QW; ++t)
                    hit |= max(max(c[t][0], c[t][1]), max(c[t][2], c[t][3])) >= V[t];
                if (ballot(hit)) {
#pragma unroll
                    for (int t = 0; t < MF_QW; ++t) {
                        // only the lanes of tiles with a kept row
                        if (max(max(c[t][0], c[t][1]), max(c[t][2], c[t][3])) >= V[t]) {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (c[t][r] >= V[t]) {
                                    // the partial distance from c': e = -c' (IP);
                                    // e = w - 2c = (w & 1) - 2c' (L2), kept here
                                    // as -2c' and completed by the select
                                    // with the row's w & 1.  Key: the value as
                                    // an order-preserving unsigned, the row id
                                    const int32_t x = MIPS ? -c[t][r] : -2 * c[t][r];
                                    const uint32_t row = (uint32_t)b0 + 16 * blk + 4 * kq + r;
                                    if (kept[t] < a.seg_cap)
                                        a.cand[segb[t] + kept[t]] =
                                            ((uint64_t)((uint32_t)x ^ 0x80000000u) << 32) | row;
                                    ++kept[t];
                                }
                        }
                    }
                }
            }
        }
    };
    using R0 = std::integral_constant<int, 0>;
    using R1 = std::integral_constant<int, 1>;
    fetch(r0, R0());
    fetch(r0 + MF_BLK, R1());
    put(0, R0());
    __syncthreads();
    // blocks in pairs: the first of a pair is in LDS buffer 0, the second in
    // buffer 1; register set 0 carries the block two ahead of buffer 0's
    for (uint64_t b0 = r0; b0 < r1; b0 += 2 * MF_BLK) {
        if (b0 + 2 * MF_BLK < r1)
            fetch(b0 + 2 * MF_BLK, R0());
        score(0, b0);
        if (b0 + MF_BLK < r1)
            put(1, R1());
        __syncthreads();
        if (b0 + MF_BLK >= r1)
            break;
        if (b0 + 3 * MF_BLK < r1)
            fetch(b0 + 3 * MF_BLK, R1());
        score(1, b0 + MF_BLK);
        if (b0 + 2 * MF_BLK < r1)
            put(0, R0());
        __syncthreads();
    }
    if (!THRESH && active) {
#pragma unroll
        for (int t = 0; t < MF_QW; ++t) {
            const uint64_t q = q0 + 16 * t + col;
            if (q < a.nq)
                a.segcnt[(q * gridDim.y + blockIdx.y) * 4 + kq] = kept[t];
        }
    }
    if (THRESH && active) {
        const uint32_t slices = gridDim.y;
#pragma unroll
        for (int t = 0; t < MF_QW; ++t) {
            const uint64_t q = q0 + 16 * t + col;
            if (q < a.nq) {
                int32_t *o = a.tv + ((q * slices + blockIdx.y) * 4 + kq) * MF_TV;
#pragma unroll
                for (int i = 0; i < MF_TV; ++i)
                    o[i] = mt[t][i];
            }
        }
    }
}

// per query (one wavefront): U = the J-th smallest of its nv values; the
// redo list's length starts at zero
__global__ __launch_bounds__(256) void mf_bound_kernel(const int32_t *__restrict__ tv, uint32_t nv,
                                                     uint32_t J, uint64_t nq,
                                                     int32_t *__restrict__ U,
                                                     uint32_t *__restrict__ redo_cnt)
{
    const uint32_t lane = threadIdx.x % WAVE;
    const uint64_t q = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *redo_cnt = 0;
    if (q >= nq)
        return;
    constexpr int PER = 4 * MF_TV * MF_MAX_TSLICES / WAVE;
    int32_t v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t i = lane + WAVE * j;
        v[j] = i < nv ? tv[q * nv + i] : INT_MAX;
    }
    int32_t cur = INT_MAX;
    for (uint32_t round = 0; round < J; ++round) {
        int32_t m = v[0];
#pragma unroll
        for (int j = 1; j < PER; ++j)
            m = min(m, v[j]);
        int32_t w = m;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            w = min(w, __shfl_xor(w, o));
        cur = w;
        // the lowest lane holding the minimum drops one copy of it
        const uint64_t holders = ballot(m == w);
        if (lane == (uint32_t)__builtin_ctzll(holders)) {
            bool done = false;
#pragma unroll
            for (int j = 0; j < PER; ++j)
                if (!done && v[j] == w) {
                    v[j] = INT_MAX;
                    done = true;
                }
        }
    }
    if (lane == 0) {
        // below the missing-row marker, so padding rows are never admitted
        U[q] = min(cur, (1 << 30) - 1);
    }
}

// rows one select workgroup sorts: 16 KB of LDS, so a CU holds 8 of them;
// the bound keeps about 4k rows per query, a query with more is redone
#define MF_SEL_MAX 2048
#define MF_SEL_SEGS 2048

// per query: gather its nseg segments, keep the rows that can be among the
// first k (a histogram cut), sort those by (partial distance, id), write the
// first k with their distances.  A query with an
// overflowed segment, fewer than k kept rows or more than the workgroup
// sorts goes on the redo list.  L2 keys arrive without the parity bit of
// their partial distance (w & 1, w = |x'|^2 from nx), added back here
// before the sort.
template <bool MIPS>
__global__ __launch_bounds__(256) void mf_select_kernel(const uint64_t *__restrict__ cand,
                                                       const uint32_t *__restrict__ segcnt,
                                                       uint32_t nseg, uint32_t seg_cap,
                                                       const uint32_t *__restrict__ nx,
                                                       const uint32_t *__restrict__ nqv,
                                                       uint32_t k, uint32_t *__restrict__ ids,
                                                       float *__restrict__ dists,
                                                       uint32_t *__restrict__ redo_idx,
                                                       uint32_t *__restrict__ redo_cnt)
{
    __shared__ uint64_t s[MF_SEL_MAX];
    __shared__ uint32_t wsum[256 / WAVE];
    __shared__ uint32_t bad;
    __shared__ uint32_t hist[256];
    __shared__ uint32_t wlo[256 / WAVE], whi[256 / WAVE], kept_n, cut_bucket;
    const uint64_t q = blockIdx.x;
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    if (tid == 0)
        bad = 0;
    // thread tid owns segments [tid * per, tid * per + per): their counts,
    // then a block-wide exclusive scan of the per-thread sums
    const uint32_t per = (nseg + 255) / 256;
    uint32_t c[MF_SEL_SEGS / 256];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t j = 0; j < MF_SEL_SEGS / 256; ++j) {
        const uint32_t i = tid * per + j;
        c[j] = j < per && i < nseg ? segcnt[q * nseg + i] : 0u;
        mine += c[j];
    }
    __syncthreads();
    bool over = false;
#pragma unroll
    for (uint32_t j = 0; j < MF_SEL_SEGS / 256; ++j)
        over |= c[j] > seg_cap;
    if (over)
        bad = 1;
    uint32_t x = mine;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o)
            x += y;
    }
    if (lane == WAVE - 1)
        wsum[wv] = x;
    __syncthreads();
    uint32_t start = x - mine, m = 0;
    for (uint32_t w = 0; w < 256 / WAVE; ++w) {
        start += w < wv ? wsum[w] : 0u;
        m += wsum[w];
    }
    if (bad || m < k || m > MF_SEL_MAX) {
        if (tid == 0)
            redo_idx[atomicAdd(redo_cnt, 1u)] = (uint32_t)q;
        return;
    }
    // each thread copies its own segments' kept rows, four loads in flight
#pragma unroll
    for (uint32_t j = 0; j < MF_SEL_SEGS / 256; ++j) {
        const uint32_t i = tid * per + j;
        if (j < per && i < nseg)
            for (uint32_t e0 = 0; e0 < c[j]; e0 += 4) {
                uint64_t v[4];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
                    if (e0 + u < c[j])
                        v[u] = cand[(q * nseg + i) * seg_cap + e0 + u];
                uint32_t par[4] = {0, 0, 0, 0};
                if (!MIPS)
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        if (e0 + u < c[j])
                            par[u] = nx[(uint32_t)v[u]] & 1u;
                // (x ^ 2^31) + 1 = (x + 1) ^ 2^31: |e| < 2^24, no carry out
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
                    if (e0 + u < c[j])
                        s[start + e0 + u] = v[u] + ((uint64_t)par[u] << 32);
            }
        start += c[j];
    }
    // keep only the rows that can be among the first k: a 256-bucket
    // histogram of the distance words, the first bucket at which the count
    // reaches k, and every row up to that bucket (rows tied with the k-th
    // share its bucket); the sort then runs on those alone
    hist[tid] = 0u;
    if (tid == 0)
        kept_n = 0u;
    __syncthreads();
    constexpr uint32_t PER = MF_SEL_MAX / 256;
    uint64_t v[PER];
    uint32_t lo = ~0u, hi = 0u;
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
        const uint32_t i = tid + 256 * u;
        v[u] = i < m ? s[i] : ~0ull;
        if (i < m) {
            lo = min(lo, (uint32_t)(v[u] >> 32));
            hi = max(hi, (uint32_t)(v[u] >> 32));
        }
    }
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    if (lane == 0) {
        wlo[wv] = lo;
        whi[wv] = hi;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t w = 0; w < 256 / WAVE; ++w) {
        lo = min(lo, wlo[w]);
        hi = max(hi, whi[w]);
    }
    uint32_t sh = 0;
    while (((hi - lo) >> sh) >= 256u)
        ++sh;
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u)
        if (tid + 256 * u < m)
            atomicAdd(&hist[((uint32_t)(v[u] >> 32) - lo) >> sh], 1u);
    __syncthreads();
    if (wv == 0) {
        // lane l holds buckets 4l .. 4l + 3
        uint32_t h4[4], t = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            h4[i] = hist[4 * lane + i];
            t += h4[i];
        }
        uint32_t x = t;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= (uint32_t)o)
                x += y;
        }
        // x - t rows lie in earlier lanes' buckets; the lane where the count
        // first reaches k names the cut bucket
        uint32_t c0 = x - t, b = 256u;
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            // counts through bucket 4 lane + i
            uint32_t through = c0;
            for (int j2 = 0; j2 <= i; ++j2)
                through += h4[j2];
            if (through >= k)
                b = 4 * lane + i;
        }
        uint32_t bm = b;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1)
            bm = min(bm, (uint32_t)__shfl_xor((int)bm, o));
        if (lane == 0)
            cut_bucket = bm;
    }
    __syncthreads();
    const uint32_t cb = cut_bucket;
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u)
        if (tid + 256 * u < m && (((uint32_t)(v[u] >> 32) - lo) >> sh) <= cb)
            s[atomicAdd(&kept_n, 1u)] = v[u];
    __syncthreads();
    const uint32_t m2 = kept_n;
    uint32_t n2 = 1;
    while (n2 < m2)
        n2 <<= 1;
    for (uint32_t i = m2 + tid; i < n2; i += 256)
        s[i] = ~0ull;
    __syncthreads();
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < n2; i += 256) {
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
    // partial distance back to the distance: + |q'|^2 (L2), - A(q) (inner product)
    const int32_t qn = (int32_t)nqv[q];
    for (uint32_t i = tid; i < k; i += 256) {
        ids[q * k + i] = key_id(s[i]);
        if (dists) {
            const int32_t e = (int32_t)((uint32_t)(s[i] >> 32) ^ 0x80000000u);
            dists[q * k + i] = (float)(MIPS ? e - qn : e + qn);
        }
    }
}

// grow-only workspace per device, reused by every call on it (calls on one
// device must not overlap in time)
struct MfWorkspace {
    uint8_t *p = nullptr;
    size_t bytes = 0;
};

static int mf_workspace(size_t bytes, uint8_t **out)
{
    static std::mutex mu;
    static MfWorkspace ws[64];
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return set_error(NBPG_EUNSUP, "device ordinal %d", dev);
    std::lock_guard<std::mutex> g(mu);
    MfWorkspace &w = ws[dev];
    if (w.bytes < bytes) {
        if (w.p) {
            HIP_TRY(hipDeviceSynchronize()); // earlier calls' work may still use it
            HIP_TRY(hipFree(w.p));
            w.p = nullptr;
            w.bytes = 0;
        }
        if (hipMalloc(&w.p, bytes) != hipSuccess)
            return set_error(NBPG_ENOMEM, "MFMA ground-truth workspace of %zu bytes", bytes);
        w.bytes = bytes;
    }
    *out = w.p;
    return NBPG_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// the scan kernel for a row width and metric, and how many of its
// workgroups one CU holds at once (registers and LDS; cached per kernel)
template <bool THRESH> static const void *mf_kernel(uint32_t KS, bool mips)
{
    switch (KS * 2 + (mips ? 1 : 0)) {
    case 2: return reinterpret_cast<const void *>(&mf_scan_kernel<1, false, THRESH>);
    case 3: return reinterpret_cast<const void *>(&mf_scan_kernel<1, true, THRESH>);
    case 4: return reinterpret_cast<const void *>(&mf_scan_kernel<2, false, THRESH>);
    case 5: return reinterpret_cast<const void *>(&mf_scan_kernel<2, true, THRESH>);
    case 6: return reinterpret_cast<const void *>(&mf_scan_kernel<3, false, THRESH>);
    case 7: return reinterpret_cast<const void *>(&mf_scan_kernel<3, true, THRESH>);
    case 8: return reinterpret_cast<const void *>(&mf_scan_kernel<4, false, THRESH>);
    default: return reinterpret_cast<const void *>(&mf_scan_kernel<4, true, THRESH>);
    }
}

static uint32_t mf_per_cu(const void *fn, uint32_t waves)
{
    static std::mutex mu;
    static std::vector<std::pair<const void *, uint32_t>> cache;
    std::lock_guard<std::mutex> g(mu);
    for (const auto &e : cache)
        if (e.first == fn)
            return e.second;
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, 64 * waves, 0) != hipSuccess || b < 1)
        b = 1;
    cache.emplace_back(fn, (uint32_t)b);
    return (uint32_t)b;
}

// slices of `rows` for `tiles` query tiles: as many workgroups as the device
// holds at once (per_cu per CU, so none waits for a second round), at
// least `min_blocks` staged blocks per slice, at most max_slices
static uint64_t mf_slice_rows(uint64_t rows, uint64_t tiles, uint32_t per_cu, uint64_t min_blocks,
                              uint64_t max_slices)
{
    uint64_t slices = std::max<uint64_t>(1, (uint64_t)per_cu * g_dev.cu / tiles);
    slices = std::min<uint64_t>(slices, std::max<uint64_t>(1, rows / (MF_BLK * min_blocks)));
    slices = std::min<uint64_t>(slices, max_slices);
    return ((rows + slices - 1) / slices + MF_BLK - 1) / MF_BLK * MF_BLK;
}

static int mf_launch(const void *fn, dim3 grid, uint32_t waves, hipStream_t s, MfArgs &a)
{
    void *args[] = {&a};
    if (hipLaunchKernel(fn, grid, dim3(64 * waves), args, 0, s) != hipSuccess)
        return set_error(NBPG_EHIP, "MFMA ground-truth scan launch");
    return NBPG_OK;
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
    const bool im = mips != 0;
    // sample stride: the bound's sample rank is about J, so about J * step
    // (4k) base rows are kept; a staged block of sample rows spans under
    // 2^31 bytes (the scan's 32-bit piece offsets)
    const uint64_t step = std::max<uint64_t>(
        1, std::min<uint64_t>({4ull * k / MF_J, n / 64, ((1ull << 31) - 1) / (MF_BLK * (uint64_t)dims)}));
    const uint64_t S = (n + step - 1) / step;
    // equal chunks of at most MF_QCHUNK queries
    const uint64_t nchunks = (nq + MF_QCHUNK - 1) / MF_QCHUNK;
    const uint64_t chunk = (nq + nchunks - 1) / nchunks;
    const uint64_t last = nq - (nq - 1) / chunk * chunk;
    const uint32_t waves = mf_waves(KS);
    const uint64_t qpb = 16ull * MF_QW * waves;
    const void *thr_fn = mf_kernel<true>(KS, im), *scan_fn = mf_kernel<false>(KS, im);
    // scan geometry of a chunk of m queries: slices of the base, and the
    // segment size for about 4k kept rows per query spread over 4 segments
    // per slice (room for three times the mean, at least 16)
    struct Geo {
        uint64_t tiles, slice_rows;
        uint32_t slices, seg_cap;
    };
    auto geo = [&](uint64_t m) {
        Geo g;
        g.tiles = (m + qpb - 1) / qpb;
        g.slice_rows = mf_slice_rows(n, g.tiles, mf_per_cu(scan_fn, waves), 4, MF_SEL_SEGS / 4);
        g.slices = (uint32_t)((n + g.slice_rows - 1) / g.slice_rows);
        const uint64_t want = std::max<uint64_t>(16, 3ull * k / g.slices + 16);
        g.seg_cap = 16;
        while (g.seg_cap < want)
            g.seg_cap <<= 1;
        return g;
    };
    const Geo gc = geo(chunk), gl = geo(last);
    const uint64_t seg_keys = std::max<uint64_t>(chunk * gc.slices * 4 * gc.seg_cap,
                                                 last * gl.slices * 4 * gl.seg_cap);
    const uint64_t segs = std::max<uint64_t>(chunk * gc.slices, last * gl.slices) * 4;
    if (seg_keys >= (1ull << 32))
        return set_error(NBPG_EUNSUP, "MFMA ground truth: k = %u needs %llu candidate slots "
                         "per query chunk (at most 2^32)", k, (unsigned long long)seg_keys);
    // workspace: nx | qn | U | redo count + list | tv | segment counts |
    // segments | the redo's partial buffer (for a full and for the last chunk)
    size_t off[9];
    off[0] = 0;
    off[1] = off[0] + align256(n * 4);
    off[2] = off[1] + align256(chunk * 4);
    off[3] = off[2] + align256(chunk * 4);
    off[4] = off[3] + align256((chunk + 1) * 4);
    off[5] = off[4] + align256(chunk * 4 * MF_TV * MF_MAX_TSLICES * 4);
    off[6] = off[5] + align256(segs * 4);
    off[7] = off[6] + align256(seg_keys * 8);
    off[8] = off[7] + align256(std::max(compgt_plain_partial_bytes(chunk, n, k, true),
                                        compgt_plain_partial_bytes(last, n, k, true)));
    uint8_t *ws = nullptr;
    NBPG_TRY(mf_workspace(off[8], &ws));
    uint32_t *nx = reinterpret_cast<uint32_t *>(ws + off[0]);
    uint32_t *nqv = reinterpret_cast<uint32_t *>(ws + off[1]);
    int32_t *U = reinterpret_cast<int32_t *>(ws + off[2]);
    uint32_t *redo_cnt = reinterpret_cast<uint32_t *>(ws + off[3]);
    uint32_t *redo_idx = redo_cnt + 1;
    int32_t *tv = reinterpret_cast<int32_t *>(ws + off[4]);
    uint32_t *segcnt = reinterpret_cast<uint32_t *>(ws + off[5]);
    uint64_t *cand = reinterpret_cast<uint64_t *>(ws + off[6]);
    uint64_t *partial = reinterpret_cast<uint64_t *>(ws + off[7]);

    hipLaunchKernelGGL(mf_norms_kernel, dim3((uint32_t)((n + 31) / 32)), dim3(256), 0, s, base, n,
                       dims, is_u8, mips, nx);
    for (uint64_t c0 = 0; c0 < nq; c0 += chunk) {
        const uint64_t m = std::min<uint64_t>(chunk, nq - c0);
        const Geo g = m == chunk ? gc : gl;
        const uint8_t *qc = queries + c0 * dims;
        uint32_t *ids = out_ids + c0 * k;
        float *dists = out_dists ? out_dists + c0 * k : nullptr;
        hipLaunchKernelGGL(mf_norms_kernel, dim3((uint32_t)((m + 31) / 32)), dim3(256), 0, s, qc,
                           m, dims, is_u8, mips, nqv);
        MfArgs a;
        memset(&a, 0, sizeof(a));
        a.base = base;
        a.queries = qc;
        a.nq = m;
        a.dims = dims;
        a.is_u8 = is_u8;
        a.nx = nx;
        a.nqv = nqv;
        a.U = U;
        a.cand = cand;
        a.segcnt = segcnt;
        a.seg_cap = g.seg_cap;
        a.tv = tv;
        // 1. bound from the sample
        a.rows = S;
        a.row_step = step;
        a.slice_rows = mf_slice_rows(S, g.tiles, mf_per_cu(thr_fn, waves), 1, MF_MAX_TSLICES);
        const uint32_t tslices = (uint32_t)((S + a.slice_rows - 1) / a.slice_rows);
        NBPG_TRY(mf_launch(thr_fn, dim3((uint32_t)g.tiles, tslices), waves, s, a));
        const uint32_t nv = 4 * MF_TV * tslices;
        hipLaunchKernelGGL(mf_bound_kernel, dim3((uint32_t)((m + 3) / 4)), dim3(256), 0, s, tv, nv,
                           std::min<uint32_t>(MF_J, nv), m, U, redo_cnt);
        // 2. scan
        a.rows = n;
        a.row_step = 1;
        a.slice_rows = g.slice_rows;
        NBPG_TRY(mf_launch(scan_fn, dim3((uint32_t)g.tiles, g.slices), waves, s, a));
        // 3. select
        if (im)
            hipLaunchKernelGGL(mf_select_kernel<true>, dim3((uint32_t)m), dim3(256), 0, s, cand,
                               segcnt, 4 * g.slices, g.seg_cap, nx, nqv, k, ids, dists,
                               redo_idx, redo_cnt);
        else
            hipLaunchKernelGGL(mf_select_kernel<false>, dim3((uint32_t)m), dim3(256), 0, s, cand,
                               segcnt, 4 * g.slices, g.seg_cap, nx, nqv, k, ids, dists,
                               redo_idx, redo_cnt);
        if (hipGetLastError() != hipSuccess)
            return set_error(NBPG_EHIP, "MFMA ground-truth launch");
        // 4. the listed queries by the plain scan
        NBPG_TRY(compgt_plain_core(dtype, base, n, dims, qc, m, dims, k, mips, ids, dists, s,
                                   partial, redo_idx, redo_cnt));
    }
    return NBPG_OK;
}

} // namespace nbpg
