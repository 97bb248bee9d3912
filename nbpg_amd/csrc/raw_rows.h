// raw_rows.h -- distances over full-precision rows, as Point::distance
// computes them for the raw PointRange of a rerank or filtering search
// (point.h:137-146 for f32: a sequential fp32 sum of squared differences;
// exact integers for u8 / i8 and the packed 4-bit types, whose two nibbles
// both unpack unsigned as (b & 0xF0) and (b << 4) & 0xF0, point.h:24-29,
// 77-120).  One lane scores one whole row; dims counts elements (bytes for
// the packed types).
#pragma once
#include "internal.h"

namespace nbpg {

template <int RT> struct raw_row;
template <> struct raw_row<NBPG_F32> {
    static __device__ __forceinline__ float dist(const uint8_t *a, const uint8_t *b, uint32_t dims)
    {
        const float *p = reinterpret_cast<const float *>(a), *q = reinterpret_cast<const float *>(b);
        float s = 0.0f;
        uint32_t i = 0;
        if ((dims & 3) == 0 && ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0) {
            // 16-byte loads, the sum still taken element by element in order
            for (; i < dims; i += 4) {
                const float4 x = *reinterpret_cast<const float4 *>(p + i);
                const float4 y = *reinterpret_cast<const float4 *>(q + i);
                float d = __fsub_rn(x.x, y.x);
                s = __fadd_rn(s, __fmul_rn(d, d));
                d = __fsub_rn(x.y, y.y);
                s = __fadd_rn(s, __fmul_rn(d, d));
                d = __fsub_rn(x.z, y.z);
                s = __fadd_rn(s, __fmul_rn(d, d));
                d = __fsub_rn(x.w, y.w);
                s = __fadd_rn(s, __fmul_rn(d, d));
            }
        }
        for (; i < dims; ++i) {
            const float d = __fsub_rn(p[i], q[i]);
            s = __fadd_rn(s, __fmul_rn(d, d));
        }
        return s;
    }
};
template <> struct raw_row<NBPG_U8> {
    static __device__ __forceinline__ float dist(const uint8_t *a, const uint8_t *b, uint32_t dims)
    {
        int32_t s = 0;
        for (uint32_t i = 0; i < dims; ++i) {
            const int32_t d = (int32_t)a[i] - (int32_t)b[i];
            s += d * d;
        }
        return (float)s;
    }
};
template <> struct raw_row<NBPG_I8> {
    static __device__ __forceinline__ float dist(const uint8_t *a, const uint8_t *b, uint32_t dims)
    {
        int32_t s = 0;
        for (uint32_t i = 0; i < dims; ++i) {
            const int32_t d = (int32_t)(int8_t)a[i] - (int32_t)(int8_t)b[i];
            s += d * d;
        }
        return (float)s;
    }
};

template <> struct raw_row<NBPG_U4X2> {
    static __device__ __forceinline__ float dist(const uint8_t *a, const uint8_t *b, uint32_t dims)
    {
        int32_t s = 0;
        for (uint32_t i = 0; i < dims; ++i) {
            const int32_t dh = (int32_t)(a[i] & 0xF0) - (int32_t)(b[i] & 0xF0);
            const int32_t dl = (int32_t)((a[i] << 4) & 0xF0) - (int32_t)((b[i] << 4) & 0xF0);
            s += dh * dh;
            s += dl * dl;
        }
        return (float)s;
    }
};

// The same distances accumulated over consecutive pieces of a row (nbytes
// each, in index order), so a row can be streamed through LDS in chunks:
// acc = raw_acc<RT>::add(acc, x, y, nbytes) ... raw_acc<RT>::fin(acc).
template <int RT> struct raw_acc;
template <> struct raw_acc<NBPG_F32> {
    typedef float T;
    static __device__ __forceinline__ T add(T s, const uint8_t *a, const uint8_t *b, uint32_t nb)
    {
        const float4 *p = reinterpret_cast<const float4 *>(a), *q = reinterpret_cast<const float4 *>(b);
        for (uint32_t i = 0; i < nb / 16; ++i) {
            const float4 x = p[i], y = q[i];
            float d = __fsub_rn(x.x, y.x);
            s = __fadd_rn(s, __fmul_rn(d, d));
            d = __fsub_rn(x.y, y.y);
            s = __fadd_rn(s, __fmul_rn(d, d));
            d = __fsub_rn(x.z, y.z);
            s = __fadd_rn(s, __fmul_rn(d, d));
            d = __fsub_rn(x.w, y.w);
            s = __fadd_rn(s, __fmul_rn(d, d));
        }
        return s;
    }
    static __device__ __forceinline__ float fin(T s) { return s; }
};
template <int RT> struct raw_acc_int {
    typedef int32_t T;
    static __device__ __forceinline__ T add(T s, const uint8_t *a, const uint8_t *b, uint32_t nb)
    {
        for (uint32_t i = 0; i < nb; ++i) {
            if (RT == NBPG_U8) {
                const int32_t d = (int32_t)a[i] - (int32_t)b[i];
                s += d * d;
            } else if (RT == NBPG_I8) {
                const int32_t d = (int32_t)(int8_t)a[i] - (int32_t)(int8_t)b[i];
                s += d * d;
            } else {
                const int32_t dh = (int32_t)(a[i] & 0xF0) - (int32_t)(b[i] & 0xF0);
                const int32_t dl = (int32_t)((a[i] << 4) & 0xF0) - (int32_t)((b[i] << 4) & 0xF0);
                s += dh * dh;
                s += dl * dl;
            }
        }
        return s;
    }
    static __device__ __forceinline__ float fin(T s) { return (float)s; }
};
template <> struct raw_acc<NBPG_U8> : raw_acc_int<NBPG_U8> {};
template <> struct raw_acc<NBPG_I8> : raw_acc_int<NBPG_I8> {};
template <> struct raw_acc<NBPG_U4X2> : raw_acc_int<NBPG_U4X2> {};

// the raw dtype chosen at run time (uniform across the wavefront)
__device__ __forceinline__ float raw_dist(int rt, const uint8_t *a, const uint8_t *b, uint32_t dims)
{
    if (rt == NBPG_F32)
        return raw_row<NBPG_F32>::dist(a, b, dims);
    if (rt == NBPG_U8)
        return raw_row<NBPG_U8>::dist(a, b, dims);
    if (rt == NBPG_I8)
        return raw_row<NBPG_I8>::dist(a, b, dims);
    return raw_row<NBPG_U4X2>::dist(a, b, dims); // U4X2 and I4X2
}

__host__ __device__ inline uint32_t raw_elem_bytes(int rt) { return rt == NBPG_F32 ? 4u : 1u; }

} // namespace nbpg
