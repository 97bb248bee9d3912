// gather_probe.hip -- roofline calibration, not part of the nbpg C-ABI.
//
// Measures the device's rate for random whole-row gathers, the access
// pattern of graph search: every lane reads `rows_per_lane` rows at
// pseudo-random positions of a table, either one lane per row (NCH dwordx4
// loads per lane) or eight lanes per 128-B row (one dwordx4 per lane, one
// instruction covering eight rows), and folds them into a checksum.
#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ uint32_t mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <int NCH>
__global__ __launch_bounds__(256) void gather_lane_per_row(const uint4 *__restrict__ table,
                                                           uint32_t nrows, uint32_t iters,
                                                           uint32_t *__restrict__ out)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t r = mix(tid * 7919u + it * 104729u) % nrows;
        const uint4 *p = table + (uint64_t)r * NCH;
        uint4 x[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            x[c] = p[c];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            acc += x[c].x ^ x[c].y ^ x[c].z ^ x[c].w;
    }
    out[tid] = acc;
}

template <int NCH>
__global__ __launch_bounds__(256) void gather_split_row(const uint4 *__restrict__ table,
                                                        uint32_t nrows, uint32_t iters,
                                                        uint32_t *__restrict__ out)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t grp = tid / NCH, c = tid % NCH;
    uint32_t acc = 0;
    for (uint32_t it = 0; it < iters; it += 8) {
        uint4 x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t r = mix(grp * 7919u + (it + u) * 104729u) % nrows;
            x[u] = table[(uint64_t)r * NCH + c];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            acc += x[u].x ^ x[u].y ^ x[u].z ^ x[u].w;
    }
    out[tid] = acc;
}

extern "C" int nbpg_probe_gather(const void *table, uint32_t nrows, uint32_t row_bytes, int split,
                                 uint32_t blocks, uint32_t iters, uint32_t *out, void *stream)
{
    if (row_bytes != 128)
        return -1;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (split)
        hipLaunchKernelGGL(gather_split_row<8>, dim3(blocks), dim3(256), 0, s,
                           (const uint4 *)table, nrows, iters, out);
    else
        hipLaunchKernelGGL(gather_lane_per_row<8>, dim3(blocks), dim3(256), 0, s,
                           (const uint4 *)table, nrows, iters, out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
