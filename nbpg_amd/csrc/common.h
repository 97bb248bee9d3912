// common.h -- device primitives shared by the gfx950 kernels.
//
// Distances are exact int32: |x-q|^2 = |x|^2 + |q|^2 - 2<x,q>, with the
// per-row norm precomputed once at upload and the dot product on the packed
// integer dot instructions (v_dot4_u32_u8, v_dot4_i32_i8, v_dot8_u32_u4).
// This equals the reference's sum of squared differences
// (point.h:35-44, 56-65, 77-90, 107-120) bit for bit, because every term is
// an integer and the final value fits in int32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nbpg.h"

#define NBPG_SENT 0xFFFFFFFFu
#define WAVE 64

namespace nbpg {

// ---------------------------------------------------------------- types
template <int DT> struct dot_traits;

template <> struct dot_traits<NBPG_U8> {
    static __device__ __forceinline__ uint32_t dot(uint32_t a, uint32_t b, uint32_t acc)
    {
        return __builtin_amdgcn_udot4(a, b, acc, false);
    }
    static constexpr uint32_t scale = 1;
};
template <> struct dot_traits<NBPG_I8> {
    static __device__ __forceinline__ uint32_t dot(uint32_t a, uint32_t b, uint32_t acc)
    {
        return (uint32_t)__builtin_amdgcn_sdot4((int)a, (int)b, (int)acc, false);
    }
    static constexpr uint32_t scale = 1;
};
// both packed 4-bit types: the reference splits a byte into b & 0xF0 and
// (b << 4) & 0xFF, i.e. 16 x (unsigned nibble); squared differences scale by 256
template <> struct dot_traits<NBPG_U4X2> {
    static __device__ __forceinline__ uint32_t dot(uint32_t a, uint32_t b, uint32_t acc)
    {
        return __builtin_amdgcn_udot8(a, b, acc, false);
    }
    static constexpr uint32_t scale = 256;
};
template <> struct dot_traits<NBPG_I4X2> : dot_traits<NBPG_U4X2> {};

// the query row as wave-uniform words: every lane holds the same query, so
// it lives in scalar registers and feeds the dot instructions' scalar operand
// the query row as plain words in every lane's vector registers: for long
// kernels whose scalar registers are already full (a spilled scalar operand
// costs a v_readlane per use)
template <int NCH>
__device__ __forceinline__ void to_lanes(const uint4 (&qv)[NCH], uint32_t (&qs)[4 * NCH])
{
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        qs[4 * c + 0] = qv[c].x;
        qs[4 * c + 1] = qv[c].y;
        qs[4 * c + 2] = qv[c].z;
        qs[4 * c + 3] = qv[c].w;
    }
}

template <int NCH>
__device__ __forceinline__ void to_uniform(const uint4 (&qv)[NCH], uint32_t (&qs)[4 * NCH])
{
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        qs[4 * c + 0] = __builtin_amdgcn_readfirstlane(qv[c].x);
        qs[4 * c + 1] = __builtin_amdgcn_readfirstlane(qv[c].y);
        qs[4 * c + 2] = __builtin_amdgcn_readfirstlane(qv[c].z);
        qs[4 * c + 3] = __builtin_amdgcn_readfirstlane(qv[c].w);
    }
}

// <x, q> over NCH 16-byte chunks of the row; q wave-uniform.  Two
// accumulators take alternate words: a dot op that reads the previous dot's
// result as its accumulator needs a wait state, two interleaved chains need
// none (integer sums: the split is exact)
template <int DT, int NCH>
__device__ __forceinline__ uint32_t dot_row(const uint4 *__restrict__ row,
                                            const uint32_t (&q)[4 * NCH])
{
    uint4 x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        x[c] = row[c];
    uint32_t a0 = 0, a1 = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        a0 = dot_traits<DT>::dot(x[c].x, q[4 * c + 0], a0);
        a1 = dot_traits<DT>::dot(x[c].y, q[4 * c + 1], a1);
        a0 = dot_traits<DT>::dot(x[c].z, q[4 * c + 2], a0);
        a1 = dot_traits<DT>::dot(x[c].w, q[4 * c + 3], a1);
    }
    return a0 + a1;
}

// exact reference distance as float: (float)(int32) sum of squares, or
// under the inner-product metric (mips != 0) the negated integer dot product
// (mips_distance, point.h:46-135)
template <int DT>
__device__ __forceinline__ float dist_from(uint32_t nx, uint32_t nq, uint32_t dot, uint32_t mips = 0)
{
    // both forms computed and one selected: a uniform branch here would split
    // the callers' dot chains apart
    const uint32_t s = nx + nq - 2u * dot; // wraps back into range: true value < 2^31
    const float f = (float)(int32_t)((mips ? dot : s) * dot_traits<DT>::scale);
    return mips ? -f : f;
}

// the four sums of one 16-byte chunk into two <x, q> and two |x|^2
// accumulators (each chain's next use is two dot ops away)
template <int DT>
__device__ __forceinline__ void dot_chunk(const uint4 &x, uint32_t q0, uint32_t q1, uint32_t q2,
                                          uint32_t q3, uint32_t (&xq)[2], uint32_t (&xx)[2])
{
    xq[0] = dot_traits<DT>::dot(x.x, q0, xq[0]);
    xx[0] = dot_traits<DT>::dot(x.x, x.x, xx[0]);
    xq[1] = dot_traits<DT>::dot(x.y, q1, xq[1]);
    xx[1] = dot_traits<DT>::dot(x.y, x.y, xx[1]);
    xq[0] = dot_traits<DT>::dot(x.z, q2, xq[0]);
    xx[0] = dot_traits<DT>::dot(x.z, x.z, xx[0]);
    xq[1] = dot_traits<DT>::dot(x.w, q3, xq[1]);
    xx[1] = dot_traits<DT>::dot(x.w, x.w, xx[1]);
}

// distance to a row without a stored norm: |x|^2 comes from the same loaded
// chunks (one more dot per word), so a distance is exactly one row gather
template <int DT, int NCH>
__device__ __forceinline__ float dist_row(const uint4 *__restrict__ row, const uint32_t (&q)[4 * NCH],
                                         uint32_t nq, uint32_t mips = 0)
{
    uint4 x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        x[c] = row[c];
    uint32_t xq[2] = {0, 0}, xx[2] = {0, 0};
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        dot_chunk<DT>(x[c], q[4 * c], q[4 * c + 1], q[4 * c + 2], q[4 * c + 3], xq, xx);
    return dist_from<DT>(xx[0] + xx[1], nq, xq[0] + xq[1], mips);
}

// same, with the query read from LDS one 16-byte chunk at a time (every lane
// reads the same address: a broadcast), so the query costs no registers
template <int DT, int NCH>
__device__ __forceinline__ float dist_row_lds(const uint4 *__restrict__ row, const uint4 *q, uint32_t nq,
                                             uint32_t mips = 0)
{
    uint4 x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        x[c] = row[c];
    uint32_t xq[2] = {0, 0}, xx[2] = {0, 0};
    // the query chunks are read one step ahead of their use, not all at once:
    // a whole query row in registers would cost 4 * NCH more of them
    uint4 qc = q[0];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const uint4 qn = q[c + 1 < NCH ? c + 1 : c];
        __builtin_amdgcn_sched_barrier(0);
        dot_chunk<DT>(x[c], qc.x, qc.y, qc.z, qc.w, xq, xx);
        qc = qn;
    }
    return dist_from<DT>(xx[0] + xx[1], nq, xq[0] + xq[1], mips);
}

template <int DT, int NCH>
__device__ __forceinline__ uint32_t norm_regs(const uint4 (&q)[NCH])
{
    // |q|^2 of a row held in vector registers (two interleaved chains)
    uint32_t a0 = 0, a1 = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        a0 = dot_traits<DT>::dot(q[c].x, q[c].x, a0);
        a1 = dot_traits<DT>::dot(q[c].y, q[c].y, a1);
        a0 = dot_traits<DT>::dot(q[c].z, q[c].z, a0);
        a1 = dot_traits<DT>::dot(q[c].w, q[c].w, a1);
    }
    return a0 + a1;
}

// ---------------------------------------------------------------- hashing
// splitmix64 finaliser (utilities.h:31-37); the visited-filter slot hash
__device__ __forceinline__ uint64_t hash64_2(uint64_t x)
{
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// ---------------------------------------------------------------- keys
// order-preserving map of a float to u32 (inner-product distances are
// negative): positive values get the top bit set, negative ones are
// inverted; -0 maps like +0, as the two compare equal
__device__ __forceinline__ uint32_t ord_f(float f)
{
    const uint32_t b = __float_as_uint(f == 0.f ? 0.f : f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord_f(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// (dist, id) records ordered as the reference's id_dist_less
// (query_params.h:115-119): (ord(dist) << 32 | id) orders as (dist asc,
// id asc) under plain u64 comparison; ~0 (EMPTY) is above every record.
__device__ __forceinline__ uint64_t make_key(float d, uint32_t id)
{
    return ((uint64_t)ord_f(d) << 32) | id;
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return (uint32_t)k; }
__device__ __forceinline__ float key_dist(uint64_t k) { return unord_f((uint32_t)(k >> 32)); }

// ---------------------------------------------------------------- wave
__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// number of set bits of mask in lanes below the calling lane
__device__ __forceinline__ uint32_t prefix_count(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, (uint32_t)__shfl_xor(v, o));
    return v;
}
// compiler + memory ordering point for lanes of one wavefront (LDS traffic
// of a wave is processed in order; this keeps the compiler from reordering)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <typename T> __device__ __forceinline__ T bcast(T v, int lane) { return __shfl(v, lane); }

// #entries e of sorted a[0:n) with e < v (lower bound), per-lane search
__device__ __forceinline__ uint32_t count_less(const uint64_t *a, uint32_t n, uint64_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
__device__ __forceinline__ bool contains(const uint64_t *a, uint32_t n, uint64_t v)
{
    uint32_t p = count_less(a, n, v);
    return p < n && a[p] == v;
}

} // namespace nbpg
