// pca.hip -- PCACompress on the device: the projection y = bias + W x
// (matmul, matrix.h:86-97), its codes (compress, matrix.h:99-127), and the
// FP64 mean / covariance a PCA fit starts from.
//
// Projection, two forms:
//   NBPG_PCA_PLAIN  one lane per (row, output), fp32 products summed in index
//                   order exactly as mips_distance (point.h:148-156) does when
//                   compiled without reassociation: y is bit-identical.
//   NBPG_PCA_BF16   v_mfma_f32_16x16x32_bf16: x and W rounded to bf16 (round
//                   to nearest even), fp32 accumulation.  A wavefront owns 32
//                   rows and all d_out outputs, streams its rows once, and
//                   quantizes in registers (the codes never round-trip HBM as
//                   floats).  The pass is HBM-bound on reading x.
// The quantizer is shared: ti = trunc(double(y + 1.0f) * 0.5 * 255), clamped
// to [0, 255], minus 128 for the signed codes, shifted right by 4 for 4-bit
// codes; chunk 2 packs byte i = code(2i+1) << 4 | code(2i) (matrix.h:104-124).
//
// Statistics: column sums and the centred Gram matrix in FP64
// (v_mfma_f64_16x16x4f64), per row slice, then a fixed-order reduction, so
// the result does not depend on scheduling.
#include <algorithm>
#include <climits>
#include <vector>

#include "internal.h"

namespace nbpg {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- inputs
// loadFloatPoint (matrix.h:27-57): u8 / i8 rows become floats
template <int XT> __device__ __forceinline__ float x_at(const void *x, uint64_t i);
template <> __device__ __forceinline__ float x_at<NBPG_F32>(const void *x, uint64_t i)
{
    return reinterpret_cast<const float *>(x)[i];
}
template <> __device__ __forceinline__ float x_at<NBPG_U8>(const void *x, uint64_t i)
{
    return (float)reinterpret_cast<const uint8_t *>(x)[i];
}
template <> __device__ __forceinline__ float x_at<NBPG_I8>(const void *x, uint64_t i)
{
    return (float)reinterpret_cast<const int8_t *>(x)[i];
}

// eight consecutive elements (i % 8 == 0, row starts 8-element aligned)
template <int XT> __device__ __forceinline__ void x_load8(const void *x, uint64_t i, float (&v)[8]);
template <> __device__ __forceinline__ void x_load8<NBPG_F32>(const void *x, uint64_t i, float (&v)[8])
{
    const float4 *p = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(x) + i);
    const float4 a = p[0], b = p[1];
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
}
template <> __device__ __forceinline__ void x_load8<NBPG_U8>(const void *x, uint64_t i, float (&v)[8])
{
    const uint2 u = *reinterpret_cast<const uint2 *>(reinterpret_cast<const uint8_t *>(x) + i);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = (float)((u.x >> (8 * j)) & 0xFF);
        v[4 + j] = (float)((u.y >> (8 * j)) & 0xFF);
    }
}
template <> __device__ __forceinline__ void x_load8<NBPG_I8>(const void *x, uint64_t i, float (&v)[8])
{
    const uint2 u = *reinterpret_cast<const uint2 *>(reinterpret_cast<const uint8_t *>(x) + i);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = (float)(int8_t)((u.x >> (8 * j)) & 0xFF);
        v[4 + j] = (float)(int8_t)((u.y >> (8 * j)) & 0xFF);
    }
}

// ---------------------------------------------------------------- quantizer
struct CodeSpec {
    uint32_t chunk; // 1 or 2 values per byte
    int delta;      // 0 (unsigned) or 128 (signed)
    int shift;      // 8 - 8 / chunk
    int mask;       // (1 << (8 / chunk)) - 1
};

inline CodeSpec code_spec(int code)
{
    CodeSpec c;
    c.chunk = (code == NBPG_CODE_U4 || code == NBPG_CODE_I4) ? 2 : 1;
    c.delta = (code == NBPG_CODE_I8 || code == NBPG_CODE_I4) ? 128 : 0;
    c.shift = 8 - (int)(8 / c.chunk);
    c.mask = (1 << (8 / c.chunk)) - 1;
    return c;
}

// double -> int as the x86-64 truncating conversion the reference compiles
// to: NaN and values outside the int range give INT_MIN
__device__ __forceinline__ int trunc_i32(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0))
        return INT_MIN;
    return (int)v;
}

__device__ __forceinline__ uint32_t code_of(float y, const CodeSpec &c)
{
    const double t = (double)(y + 1.0f) * 0.5 * 255.0;
    int ti = trunc_i32(t);
    ti = ti < 0 ? 0 : (ti > 255 ? 255 : ti);
    ti = (ti - c.delta) >> c.shift;
    return (uint32_t)(ti & c.mask);
}

__global__ void pca_quantize_kernel(const float *__restrict__ y, uint64_t n, uint32_t d, CodeSpec c,
                                    uint8_t *__restrict__ out)
{
    const uint32_t od = d / c.chunk;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * od)
        return;
    const uint64_t p = t / od, i = t % od;
    const float *v = y + p * d;
    uint32_t r = 0;
    for (uint32_t j = 0; j < c.chunk; ++j) {
        r = (r << (8 / c.chunk)) & 0xFF;
        r = (r + code_of(v[(i + 1) * c.chunk - 1 - j], c)) & 0xFF;
    }
    out[t] = (uint8_t)r;
}

// ---------------------------------------------------------------- plain
__global__ void pca_transpose_kernel(const float *__restrict__ W, uint32_t d_out, uint32_t d_in,
                                     float *__restrict__ Wt)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)d_out * d_in)
        return;
    const uint32_t o = (uint32_t)(t / d_in), i = (uint32_t)(t % d_in);
    Wt[(uint64_t)i * d_out + o] = W[t];
}

// lane = (row, output); the lanes of a row read x[row][i] together and
// W^T[i][output] coalesced
template <int XT>
__global__ __launch_bounds__(256) void pca_plain_kernel(const void *__restrict__ x, uint64_t n,
                                                        uint32_t d_in, const float *__restrict__ Wt,
                                                        const float *__restrict__ bias,
                                                        uint32_t d_out, float *__restrict__ y)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * d_out)
        return;
    const uint64_t row = t / d_out;
    const uint32_t o = (uint32_t)(t % d_out);
    float s = 0.0f;
    for (uint32_t i = 0; i < d_in; ++i)
        s = __fadd_rn(s, __fmul_rn(Wt[(uint64_t)i * d_out + o], x_at<XT>(x, row * d_in + i)));
    y[t] = __fsub_rn(bias[o], -s); // bias[jj] - mips_distance (matrix.h:94)
}

// ---------------------------------------------------------------- bf16 MFMA
// W as bf16, rows zero-padded to kp (a multiple of 32) columns
__global__ void pca_w_bf16_kernel(const float *__restrict__ W, uint32_t d_out, uint32_t d_in,
                                  uint32_t kp, __bf16 *__restrict__ Wb)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)d_out * kp)
        return;
    const uint32_t o = (uint32_t)(t / kp), i = (uint32_t)(t % kp);
    Wb[t] = (__bf16)(i < d_in ? W[(uint64_t)o * d_in + i] : 0.0f);
}

struct PcaArgs {
    const void *x;
    uint64_t n;
    uint32_t d_in, kp, d_out;
    const __bf16 *Wb;
    const float *bias;
    CodeSpec c;
    int with_codes;
    uint8_t *codes;
    float *y;
};

#define PCA_TILES 4                // 16-row MFMA tiles per wavefront
#define PCA_ROWS (16 * PCA_TILES) // rows per wavefront: W is read once per PCA_ROWS rows

// One wavefront's PCA_ROWS rows: acc = X_rows . W^T on bf16 MFMA, then bias,
// y and codes.  W fragments come from Wp (row pitch wpitch elements), either
// global memory or the workgroup's LDS copy.
template <int XT, int NT, typename WPtr>
__device__ __forceinline__ void pca_bf16_rows(const PcaArgs &a, uint64_t row0, WPtr Wp, uint32_t wpitch,
                                              uint32_t lane)
{
    const uint32_t r16 = lane % 16, kg = lane / 16;
    v4f acc[PCA_TILES][NT];
#pragma unroll
    for (int t = 0; t < PCA_TILES; ++t)
#pragma unroll
        for (int j = 0; j < NT; ++j)
            acc[t][j] = v4f{0.f, 0.f, 0.f, 0.f};
    for (uint32_t k0 = 0; k0 < a.kp; k0 += 32) {
        const uint32_t kk = k0 + 8 * kg;
        v8bf A[PCA_TILES];
#pragma unroll
        for (int t = 0; t < PCA_TILES; ++t) {
            const uint64_t row = row0 + 16 * t + r16;
            float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (row < a.n && kk < a.d_in)
                x_load8<XT>(a.x, row * a.d_in + kk, v);
#pragma unroll
            for (int e = 0; e < 8; ++e)
                A[t][e] = (__bf16)v[e];
        }
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const v8bf B = *reinterpret_cast<const v8bf *>(Wp + (uint64_t)(16 * j + r16) * wpitch + kk);
#pragma unroll
            for (int t = 0; t < PCA_TILES; ++t)
                acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t], B, acc[t][j], 0, 0, 0);
        }
    }
    // lane holds outputs (row0 + 16t + 4kg + r, 16j + r16)
    const uint32_t od = a.d_out / a.c.chunk;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const uint32_t col = 16 * j + r16;
        const float b = a.bias[col];
#pragma unroll
        for (int t = 0; t < PCA_TILES; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint64_t row = row0 + 16 * t + 4 * kg + r;
                const float yv = b + acc[t][j][r];
                if (a.y && row < a.n)
                    a.y[row * a.d_out + col] = yv;
                if (a.with_codes) {
                    const uint32_t cv = code_of(yv, a.c);
                    if (a.c.chunk == 1) {
                        if (row < a.n)
                            a.codes[row * od + col] = (uint8_t)cv;
                    } else {
                        const uint32_t hi = __shfl_down(cv, 1); // column col + 1
                        if ((r16 & 1) == 0 && row < a.n)
                            a.codes[row * od + col / 2] = (uint8_t)(((hi << 4) + cv) & 0xFF);
                    }
                }
            }
    }
}

// W read from global memory (through L1/L2) by every wavefront
template <int XT, int NT>
__global__ __launch_bounds__(256) void pca_bf16_kernel(PcaArgs a)
{
    const uint32_t lane = threadIdx.x % WAVE;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint64_t row0 = wave * PCA_ROWS;
    if (row0 >= a.n)
        return;
    pca_bf16_rows<XT, NT>(a, row0, a.Wb, a.kp, lane);
}

// LDS variant: a workgroup of PCA_LDS_WAVES wavefronts copies W (16 NT rows
// of kp bf16) into LDS, with a 16-byte pad per row against bank conflicts,
// and each of its wavefronts then projects one row block reading W
// fragments from LDS (one copy of W per PCA_LDS_WAVES blocks instead of one
// read per block).
#define PCA_LDS_WAVES 16
template <int XT, int NT>
__global__ __launch_bounds__(64 * PCA_LDS_WAVES) void pca_bf16_lds_kernel(PcaArgs a)
{
    extern __shared__ __align__(16) uint8_t lds_w[];
    const uint32_t wpitch = a.kp + 8; // elements
    const uint32_t row16s = a.kp / 8;  // 16-byte pieces per W row
    for (uint32_t i = threadIdx.x; i < 16 * NT * row16s; i += blockDim.x) {
        const uint32_t r = i / row16s, c = i % row16s;
        *reinterpret_cast<uint4 *>(lds_w + ((uint64_t)r * wpitch + 8 * c) * 2) =
            *reinterpret_cast<const uint4 *>(a.Wb + (uint64_t)r * a.kp + 8 * c);
    }
    __syncthreads();
    const __bf16 *Wl = reinterpret_cast<const __bf16 *>(lds_w);
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const uint64_t row0 = ((uint64_t)blockIdx.x * (blockDim.x / WAVE) + wv) * PCA_ROWS;
    if (row0 < a.n)
        pca_bf16_rows<XT, NT>(a, row0, Wl, wpitch, lane);
}

// ---------------------------------------------------------------- FP64 stats
#define ST_SLICES 64 // fewest row slices of the column sums

template <int XT>
__global__ void pca_colsum_kernel(const void *__restrict__ x, uint64_t n, uint32_t d,
                                  uint64_t slice_rows, double *__restrict__ part)
{
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= d)
        return;
    const uint64_t r0 = (uint64_t)blockIdx.y * slice_rows, r1 = min(n, r0 + slice_rows);
    double s = 0.0;
    for (uint64_t r = r0; r < r1; ++r)
        s += (double)x_at<XT>(x, r * d + col);
    part[(uint64_t)blockIdx.y * d + col] = s;
}

// column means from the slices' sums: 64 columns per workgroup, 4 threads
// per column each summing every 4th slice, then the 4 partials in a fixed
// order (deterministic)
__global__ __launch_bounds__(256) void pca_mean_kernel(const double *__restrict__ part,
                                                       uint32_t slices, uint32_t d, uint64_t n,
                                                       double *__restrict__ mean)
{
    __shared__ double ps[4][64];
    const uint32_t c = threadIdx.x % 64, g = threadIdx.x / 64;
    const uint32_t col = blockIdx.x * 64 + c;
    double s = 0.0;
    if (col < d)
        for (uint32_t i = g; i < slices; i += 4)
            s += part[(uint64_t)i * d + col];
    ps[g][c] = s;
    __syncthreads();
    if (g == 0 && col < d)
        mean[col] = (((ps[0][c] + ps[1][c]) + ps[2][c]) + ps[3][c]) / (double)n;
}

// one 64 x 64 tile (ti <= tj) of sum_r (x_r - mean)(x_r - mean)^T over a row
// slice; wave w owns tile rows 16w..16w+15.  Slabs of GR_K rows of both
// column blocks are loaded once per workgroup (coalesced, the next slab in
// registers while this one is summed), centred in FP64 and staged in LDS;
// the MFMAs then run over the slab 4 rows at a time, in row order, so every
// entry's sum is formed in the same order as a plain 4-rows-per-step loop.
#define GR_K 32     // rows per slab
#define GR_PITCH 66 // LDS row pitch in doubles (pad against bank conflicts)

template <int XT>
__global__ __launch_bounds__(256) void pca_gram_kernel(const void *__restrict__ x, uint64_t n,
                                                       uint32_t d, const double *__restrict__ mean,
                                                       const uint32_t *__restrict__ pairs,
                                                       uint64_t slice_rows,
                                                       double *__restrict__ part)
{
    __shared__ double sa[GR_K * GR_PITCH], sb[GR_K * GR_PITCH];
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    const uint32_t r16 = lane % 16, kg = lane / 16;
    const uint32_t pr = pairs[blockIdx.x];
    const uint32_t ti = pr >> 16, tj = pr & 0xFFFF;
    const bool diag = ti == tj;
    const double *sbr = diag ? sa : sb;
    // this thread's elements of a slab: rows 4j + tid / 64 (j < GR_K / 4), column tid % 64
    const uint32_t col = tid % 64, roff = tid / 64;
    const uint32_t ca = ti * 64 + col, cb = tj * 64 + col;
    const double ma = ca < d ? mean[ca] : 0.0, mb = cb < d ? mean[cb] : 0.0;
    const uint64_t r0 = (uint64_t)blockIdx.y * slice_rows, r1 = min(n, r0 + slice_rows);
    constexpr int J = GR_K / 4;
    float fa[J], fb[J];
    auto fetch = [&](uint64_t k0) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const uint64_t row = k0 + 4 * j + roff;
            fa[j] = (row < r1 && ca < d) ? x_at<XT>(x, row * d + ca) : 0.0f;
            fb[j] = (!diag && row < r1 && cb < d) ? x_at<XT>(x, row * d + cb) : 0.0f;
        }
    };
    auto put = [&](uint64_t k0) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const uint64_t row = k0 + 4 * j + roff;
            const uint32_t at = (4 * j + roff) * GR_PITCH + col;
            sa[at] = (row < r1 && ca < d) ? (double)fa[j] - ma : 0.0;
            if (!diag)
                sb[at] = (row < r1 && cb < d) ? (double)fb[j] - mb : 0.0;
        }
    };
    v4d acc[4];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
        acc[jj] = v4d{0.0, 0.0, 0.0, 0.0};
    fetch(r0);
    for (uint64_t k0 = r0; k0 < r1; k0 += GR_K) {
        put(k0);
        __syncthreads();
        if (k0 + GR_K < r1)
            fetch(k0 + GR_K); // in flight while this slab is summed
#pragma unroll
        for (int kk = 0; kk < GR_K; kk += 4) {
            const double av = sa[(kk + kg) * GR_PITCH + 16 * wv + r16];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const double bv = sbr[(kk + kg) * GR_PITCH + 16 * jj + r16];
                acc[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[jj], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // the f64 16x16x4 result interleaves rows: register r of lane l holds
    // row 4r + l/16, column l%16 (unlike the 4(l/16) + r of the i8 / bf16
    // forms)
    double *out = part + ((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * 4096;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            out[(16 * wv + 4 * r + kg) * 64 + 16 * jj + r16] = acc[jj][r];
}

__global__ void pca_cov_reduce_kernel(const double *__restrict__ part, const uint32_t *__restrict__ pairs,
                                      uint32_t npairs, uint32_t slices, uint32_t d, uint64_t n,
                                      double *__restrict__ cov)
{
    const uint32_t p = blockIdx.x, e = blockIdx.y * blockDim.x + threadIdx.x; // e < 4096
    const uint32_t ti = pairs[p] >> 16, tj = pairs[p] & 0xFFFF;
    const uint32_t i = ti * 64 + e / 64, j = tj * 64 + e % 64;
    // a diagonal tile writes its upper half only: one writer per entry
    if (i >= d || j >= d || (ti == tj && e / 64 > e % 64))
        return;
    double s = 0.0;
    for (uint32_t sl = 0; sl < slices; ++sl)
        s += part[((uint64_t)sl * npairs + p) * 4096 + e];
    s /= (double)(n - 1);
    cov[(uint64_t)i * d + j] = s;
    cov[(uint64_t)j * d + i] = s;
}

} // namespace nbpg

using namespace nbpg;

static bool x_dtype_ok(int t) { return t == NBPG_F32 || t == NBPG_U8 || t == NBPG_I8; }

template <typename F> static void dispatch_x(int t, F &&f)
{
    if (t == NBPG_F32)
        f(std::integral_constant<int, NBPG_F32>());
    else if (t == NBPG_U8)
        f(std::integral_constant<int, NBPG_U8>());
    else
        f(std::integral_constant<int, NBPG_I8>());
}

extern "C" int nbpg_pca_quantize(const float *d_y, uint64_t n, uint32_t d, int code, uint8_t *d_out,
                                 void *stream)
{
    NBPG_TRY(ensure_init());
    if (code < NBPG_CODE_U8 || code > NBPG_CODE_I4 || (n && (!d_y || !d_out)))
        return set_error(NBPG_EINVAL, "bad argument");
    const CodeSpec c = code_spec(code);
    if (d % c.chunk)
        return set_error(NBPG_EINVAL, "4-bit codes need an even dimension (%u)", d);
    const uint64_t total = n * (d / c.chunk);
    if (total == 0)
        return NBPG_OK;
    hipLaunchKernelGGL(pca_quantize_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0,
                       as_stream(stream), d_y, n, d, c, d_out);
    return hipGetLastError() == hipSuccess ? NBPG_OK : set_error(NBPG_EHIP, "quantize launch");
}

extern "C" int nbpg_pca_project(int x_dtype, const void *d_x, uint64_t n, uint32_t d_in,
                                const float *d_W, const float *d_bias, uint32_t d_out, int mode,
                                int code, uint8_t *d_out_codes, float *d_out_y, void *stream)
{
    NBPG_TRY(ensure_init());
    if (!x_dtype_ok(x_dtype) || code < NBPG_CODE_F32 || code > NBPG_CODE_I4 || d_in == 0 ||
        d_out == 0)
        return set_error(NBPG_EINVAL, "bad argument");
    if (n && (!d_x || !d_W || !d_bias || (code != NBPG_CODE_F32 && !d_out_codes) ||
              (code == NBPG_CODE_F32 && !d_out_y)))
        return set_error(NBPG_EINVAL, "null buffer");
    const CodeSpec c = code_spec(code == NBPG_CODE_F32 ? NBPG_CODE_U8 : code);
    if (code != NBPG_CODE_F32 && d_out % c.chunk)
        return set_error(NBPG_EINVAL, "4-bit codes need an even output dimension (%u)", d_out);
    if (n == 0)
        return NBPG_OK;
    hipStream_t s = as_stream(stream);
    if (mode == NBPG_PCA_PLAIN) {
        float *Wt = nullptr, *y = d_out_y;
        if (hipMallocAsync((void **)&Wt, (size_t)d_in * d_out * 4, s) != hipSuccess ||
            (!y && hipMallocAsync((void **)&y, n * d_out * 4, s) != hipSuccess)) {
            hipFreeAsync(Wt, s);
            return set_error(NBPG_ENOMEM, "projection workspace");
        }
        const uint64_t wn = (uint64_t)d_in * d_out, tn = n * d_out;
        hipLaunchKernelGGL(pca_transpose_kernel, dim3((uint32_t)((wn + 255) / 256)), dim3(256), 0,
                           s, d_W, d_out, d_in, Wt);
        dispatch_x(x_dtype, [&](auto XT) {
            hipLaunchKernelGGL((pca_plain_kernel<XT.value>), dim3((uint32_t)((tn + 255) / 256)),
                               dim3(256), 0, s, d_x, n, d_in, (const float *)Wt, d_bias, d_out, y);
        });
        int rc = hipGetLastError() == hipSuccess ? NBPG_OK : set_error(NBPG_EHIP, "projection launch");
        if (rc == NBPG_OK && code != NBPG_CODE_F32)
            rc = nbpg_pca_quantize(y, n, d_out, code, d_out_codes, stream);
        hipFreeAsync(Wt, s);
        if (y != d_out_y)
            hipFreeAsync(y, s);
        return rc;
    }
    if (mode != NBPG_PCA_BF16)
        return set_error(NBPG_EINVAL, "projection mode %d", mode);
    if (d_in % 8 || d_out % 16 || d_out > 128 || (reinterpret_cast<uintptr_t>(d_x) & 15))
        return set_error(NBPG_EUNSUP,
                         "bf16 projection: d_in %% 8 == 0, d_out a multiple of 16 up to 128, "
                         "16-byte aligned rows (d_in %u, d_out %u)", d_in, d_out);
    const uint32_t kp = (d_in + 31) / 32 * 32;
    __bf16 *Wb = nullptr;
    if (hipMallocAsync((void **)&Wb, (size_t)d_out * kp * 2, s) != hipSuccess)
        return set_error(NBPG_ENOMEM, "projection workspace");
    const uint64_t wn = (uint64_t)d_out * kp;
    hipLaunchKernelGGL(pca_w_bf16_kernel, dim3((uint32_t)((wn + 255) / 256)), dim3(256), 0, s, d_W,
                       d_out, d_in, kp, Wb);
    PcaArgs a;
    a.x = d_x;
    a.n = n;
    a.d_in = d_in;
    a.kp = kp;
    a.d_out = d_out;
    a.Wb = Wb;
    a.bias = d_bias;
    a.c = c;
    a.with_codes = code != NBPG_CODE_F32;
    a.codes = d_out_codes;
    a.y = d_out_y;
    const uint64_t waves = (n + PCA_ROWS - 1) / PCA_ROWS;
    const dim3 grid((uint32_t)((waves + 3) / 4)), block(256);
    // W in LDS when it fits (64 x 960 bf16: 124 KB) and the rows fill more
    // than one workgroup per CU
    const uint64_t lds_bytes = (uint64_t)d_out * (kp + 8) * 2;
    const bool in_lds = lds_bytes <= (uint64_t)g_dev.lds_max && waves >= 2ull * g_dev.cu * PCA_LDS_WAVES;
    const dim3 lgrid((uint32_t)((waves + PCA_LDS_WAVES - 1) / PCA_LDS_WAVES)), lblock(64 * PCA_LDS_WAVES);
    dispatch_x(x_dtype, [&](auto XT) {
        constexpr int X = XT.value;
#define NBPG_PCA_LAUNCH(T)                                                                         \
    do {                                                                                           \
        if (in_lds)                                                                                \
            hipLaunchKernelGGL((pca_bf16_lds_kernel<X, T>), lgrid, lblock, (uint32_t)lds_bytes, s, a); \
        else                                                                                       \
            hipLaunchKernelGGL((pca_bf16_kernel<X, T>), grid, block, 0, s, a);                     \
    } while (0)
        switch (d_out / 16) {
        case 1: NBPG_PCA_LAUNCH(1); break;
        case 2: NBPG_PCA_LAUNCH(2); break;
        case 3: NBPG_PCA_LAUNCH(3); break;
        case 4: NBPG_PCA_LAUNCH(4); break;
        case 5: NBPG_PCA_LAUNCH(5); break;
        case 6: NBPG_PCA_LAUNCH(6); break;
        case 7: NBPG_PCA_LAUNCH(7); break;
        default: NBPG_PCA_LAUNCH(8); break;
        }
#undef NBPG_PCA_LAUNCH
    });
    const int rc = hipGetLastError() == hipSuccess ? NBPG_OK : set_error(NBPG_EHIP, "bf16 projection launch");
    hipFreeAsync(Wb, s);
    return rc;
}

extern "C" int nbpg_pca_stats_f64(int x_dtype, const void *d_x, uint64_t n, uint32_t d,
                                  double *d_mean, double *d_cov, void *stream)
{
    NBPG_TRY(ensure_init());
    if (!x_dtype_ok(x_dtype) || !d_x || !d_mean || n < 2 || d == 0 || d > 65535 * 64)
        return set_error(NBPG_EINVAL, "bad argument (n >= 2 rows needed)");
    hipStream_t s = as_stream(stream);
    // row slices: enough workgroups to keep every CU streaming (each thread
    // walks its column down one slice), at least 256 rows each
    const uint64_t col_blocks = (d + 255) / 256;
    uint64_t want = std::max<uint64_t>(ST_SLICES, (uint64_t)g_dev.cu * 16 / col_blocks);
    want = std::min<uint64_t>(want, std::max<uint64_t>(1, n / 256));
    const uint64_t slice_rows = (n + want - 1) / want;
    const uint32_t slices = (uint32_t)((n + slice_rows - 1) / slice_rows);
    double *part = nullptr;
    if (hipMallocAsync((void **)&part, (size_t)slices * d * 8, s) != hipSuccess)
        return set_error(NBPG_ENOMEM, "statistics workspace");
    dispatch_x(x_dtype, [&](auto XT) {
        hipLaunchKernelGGL((pca_colsum_kernel<XT.value>), dim3((d + 255) / 256, slices), dim3(256),
                           0, s, d_x, n, d, slice_rows, part);
    });
    hipLaunchKernelGGL(pca_mean_kernel, dim3((d + 63) / 64), dim3(256), 0, s, part, slices, d, n,
                       d_mean);
    hipFreeAsync(part, s);
    if (hipGetLastError() != hipSuccess)
        return set_error(NBPG_EHIP, "mean launch");
    if (!d_cov)
        return NBPG_OK;
    // upper-triangle 64 x 64 tile pairs
    const uint32_t T = (d + 63) / 64, npairs = T * (T + 1) / 2;
    std::vector<uint32_t> hp;
    for (uint32_t i = 0; i < T; ++i)
        for (uint32_t j = i; j < T; ++j)
            hp.push_back(i << 16 | j);
    // slices: enough workgroups to fill the device, at least 4096 rows each
    uint64_t gs = std::max<uint64_t>(1, (uint64_t)g_dev.cu * 8 / npairs);
    gs = std::min<uint64_t>(gs, std::max<uint64_t>(1, n / 4096));
    const uint64_t grow = ((n + gs - 1) / gs + 3) / 4 * 4;
    gs = (n + grow - 1) / grow;
    uint32_t *pairs = nullptr;
    if (hipMallocAsync((void **)&pairs, npairs * 4ull, s) != hipSuccess ||
        hipMallocAsync((void **)&part, gs * npairs * 4096ull * 8, s) != hipSuccess) {
        hipFreeAsync(pairs, s);
        return set_error(NBPG_ENOMEM, "covariance workspace");
    }
    int rc = NBPG_OK;
    if (hipMemcpyAsync(pairs, hp.data(), npairs * 4ull, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = set_error(NBPG_EHIP, "pair upload");
    if (rc == NBPG_OK) {
        dispatch_x(x_dtype, [&](auto XT) {
            hipLaunchKernelGGL((pca_gram_kernel<XT.value>), dim3(npairs, (uint32_t)gs), dim3(256),
                               0, s, d_x, n, d, (const double *)d_mean, (const uint32_t *)pairs,
                               grow, part);
        });
        hipLaunchKernelGGL(pca_cov_reduce_kernel, dim3(npairs, 4096 / 256), dim3(256), 0, s, part,
                           pairs, npairs, (uint32_t)gs, d, n, d_cov);
        if (hipGetLastError() != hipSuccess)
            rc = set_error(NBPG_EHIP, "covariance launch");
        // hp must outlive the upload
        if (hipStreamSynchronize(s) != hipSuccess && rc == NBPG_OK)
            rc = set_error(NBPG_EHIP, "covariance");
    }
    hipFreeAsync(pairs, s);
    hipFreeAsync(part, s);
    return rc;
}
