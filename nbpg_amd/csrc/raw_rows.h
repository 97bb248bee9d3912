// raw_rows.h -- distances over full-precision rows, as Point::distance
// computes them for the raw PointRange of a rerank or filtering search.
// Squared L2 (euclidean_distance, point.h:35-146) or, with IP, the negated
// inner product (mips_distance, point.h:46-156).  f32 rows are summed
// element by element in index order in fp32 (products and differences
// rounded on their own, no fused multiply-add), so a float distance is the
// source-order sum bit for bit; integer rows are exact.  The packed 4-bit
// types unpack both nibbles unsigned as (b & 0xF0) and (b << 4) & 0xF0
// (point.h:24-29, 77-135).  dims counts elements (bytes for the packed types).
#pragma once
#include "internal.h"

namespace nbpg {

// one element's term of the sum
template <int RT, bool IP> struct raw_term;
template <bool IP> struct raw_term<NBPG_F32, IP> {
    static __device__ __forceinline__ float add(float s, float x, float y)
    {
        if (IP)
            return __fadd_rn(s, __fmul_rn(x, y));
        const float d = __fsub_rn(x, y);
        return __fadd_rn(s, __fmul_rn(d, d));
    }
};
template <bool IP> struct raw_term<NBPG_U8, IP> {
    static __device__ __forceinline__ int32_t add(int32_t s, uint8_t x, uint8_t y)
    {
        const int32_t a = x, b = y;
        return IP ? s + a * b : s + (a - b) * (a - b);
    }
};
template <bool IP> struct raw_term<NBPG_I8, IP> {
    static __device__ __forceinline__ int32_t add(int32_t s, uint8_t x, uint8_t y)
    {
        const int32_t a = (int8_t)x, b = (int8_t)y;
        return IP ? s + a * b : s + (a - b) * (a - b);
    }
};
template <bool IP> struct raw_term<NBPG_U4X2, IP> {
    static __device__ __forceinline__ int32_t add(int32_t s, uint8_t x, uint8_t y)
    {
        const int32_t xh = x & 0xF0, yh = y & 0xF0, xl = (x << 4) & 0xF0, yl = (y << 4) & 0xF0;
        if (IP)
            return s + xh * yh + xl * yl;
        return s + (xh - yh) * (xh - yh) + (xl - yl) * (xl - yl);
    }
};

// the same sum over consecutive pieces of a row (nbytes each, in index
// order), so a row can be streamed through LDS in chunks:
// acc = raw_acc<RT, IP>::add(acc, x, y, nbytes) ... raw_acc<RT, IP>::fin(acc)
template <int RT, bool IP> struct raw_acc {
    typedef int32_t T;
    static __device__ __forceinline__ T add(T s, const uint8_t *a, const uint8_t *b, uint32_t nb)
    {
        for (uint32_t i = 0; i < nb; ++i)
            s = raw_term<RT, IP>::add(s, a[i], b[i]);
        return s;
    }
    static __device__ __forceinline__ float fin(T s) { return IP ? -(float)s : (float)s; }
};
template <bool IP> struct raw_acc<NBPG_F32, IP> {
    typedef float T;
    static __device__ __forceinline__ T add(T s, const uint8_t *a, const uint8_t *b, uint32_t nb)
    {
        const float4 *p = reinterpret_cast<const float4 *>(a), *q = reinterpret_cast<const float4 *>(b);
        for (uint32_t i = 0; i < nb / 16; ++i) {
            const float4 x = p[i], y = q[i];
            s = raw_term<NBPG_F32, IP>::add(s, x.x, y.x);
            s = raw_term<NBPG_F32, IP>::add(s, x.y, y.y);
            s = raw_term<NBPG_F32, IP>::add(s, x.z, y.z);
            s = raw_term<NBPG_F32, IP>::add(s, x.w, y.w);
        }
        return s;
    }
    static __device__ __forceinline__ float fin(T s) { return IP ? -s : s; }
};

// one whole row by one lane
template <int RT, bool IP> struct raw_row {
    static __device__ __forceinline__ float dist(const uint8_t *a, const uint8_t *b, uint32_t dims)
    {
        int32_t s = 0;
        for (uint32_t i = 0; i < dims; ++i)
            s = raw_term<RT, IP>::add(s, a[i], b[i]);
        return IP ? -(float)s : (float)s;
    }
};
template <bool IP> struct raw_row<NBPG_F32, IP> {
    static __device__ __forceinline__ float dist(const uint8_t *a, const uint8_t *b, uint32_t dims)
    {
        const float *p = reinterpret_cast<const float *>(a), *q = reinterpret_cast<const float *>(b);
        float s = 0.0f;
        uint32_t i = 0;
        if ((dims & 3) == 0 && ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0)
            s = raw_acc<NBPG_F32, IP>::add(s, a, b, dims * 4), i = dims; // 16-byte loads
        for (; i < dims; ++i)
            s = raw_term<NBPG_F32, IP>::add(s, p[i], q[i]);
        return IP ? -s : s;
    }
};

// the raw dtype and metric chosen at run time (uniform across the wavefront)
template <bool IP>
__device__ __forceinline__ float raw_dist_t(int rt, const uint8_t *a, const uint8_t *b, uint32_t dims)
{
    if (rt == NBPG_F32)
        return raw_row<NBPG_F32, IP>::dist(a, b, dims);
    if (rt == NBPG_U8)
        return raw_row<NBPG_U8, IP>::dist(a, b, dims);
    if (rt == NBPG_I8)
        return raw_row<NBPG_I8, IP>::dist(a, b, dims);
    return raw_row<NBPG_U4X2, IP>::dist(a, b, dims); // U4X2 and I4X2
}
__device__ __forceinline__ float raw_dist(int rt, const uint8_t *a, const uint8_t *b, uint32_t dims,
                                          uint32_t mips)
{
    return mips ? raw_dist_t<true>(rt, a, b, dims) : raw_dist_t<false>(rt, a, b, dims);
}

__host__ __device__ inline uint32_t raw_elem_bytes(int rt) { return rt == NBPG_F32 ? 4u : 1u; }

} // namespace nbpg
