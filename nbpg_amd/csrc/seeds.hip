// seeds.hip -- start nodes for the FAST search: Lloyd k-means over a sample
// of the index's rows, then the row nearest each centre.
//
// The reference hands its search a list of start points through get_sp
// (check_nn_recall.h:282-297; query_params.h:206-210 seeds the beam with
// them); which points to use is the caller's choice.  This picks `count`
// of them spread over the data, on the device, deterministically:
//   * sample: rows splitmix64(seed, i) mod n, i < sample;
//   * vectors as the distance sees them (u8 / i8 values; the two nibbles of
//     a packed byte as 16-step values, point.h:27-32);
//   * centres start at the first `count` sample rows; each Lloyd step
//     assigns a row to its nearest centre (FP64 distances, lowest index on
//     ties) and sums the rows per centre in exact integers (so the result
//     does not depend on the order of the atomics), then divides in FP64;
//   * the nearest sample row to each centre (lowest sample index on ties);
//     the distinct ids, ascending, are the seeds.
#include <algorithm>
#include <vector>

#include "internal.h"

namespace nbpg {

#define SEED_MAXK 256 // centres
#define SEED_MAXD 256 // values per row as the distance sees them

__device__ __forceinline__ uint64_t seed_mix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// value j of row r as the distance sees it (D = logical values per row)
__device__ __forceinline__ int seed_value(const uint8_t *row, int dtype, uint32_t j)
{
    if (dtype == NBPG_U4X2 || dtype == NBPG_I4X2) {
        const uint32_t b = row[j / 2];
        return (j & 1) ? (int)((b << 4) & 0xFF) : (int)(b & 0xF0);
    }
    return dtype == NBPG_I8 ? (int)(int8_t)row[j] : (int)row[j];
}

// sample row ids
__global__ void seed_sample_kernel(uint64_t n, uint32_t S, uint64_t seed, uint32_t *__restrict__ ids)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S)
        ids[i] = (uint32_t)(seed_mix(seed ^ ((uint64_t)i << 20) ^ i) % n);
}

// centres c <- sample rows 0..k-1
__global__ void seed_init_kernel(const uint8_t *__restrict__ vecs, uint32_t stride, int dtype,
                                 const uint32_t *__restrict__ ids, uint32_t k, uint32_t D,
                                 double *__restrict__ c)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < k * D)
        c[t] = (double)seed_value(vecs + (uint64_t)ids[t / D] * stride, dtype, t % D);
}

// one Lloyd assignment: label per sample row, integer sums per centre
// (one thread per sample row; the centres in LDS)
__global__ __launch_bounds__(256) void seed_assign_kernel(
    const uint8_t *__restrict__ vecs, uint32_t stride, int dtype, const uint32_t *__restrict__ ids,
    uint32_t S, const double *__restrict__ c, uint32_t k, uint32_t D, uint32_t *__restrict__ label,
    long long *__restrict__ sums, uint32_t *__restrict__ counts)
{
    extern __shared__ double cl[];
    for (uint32_t t = threadIdx.x; t < k * D; t += blockDim.x)
        cl[t] = c[t];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S)
        return;
    const uint8_t *row = vecs + (uint64_t)ids[i] * stride;
    double best = 0.0;
    uint32_t bi = 0;
    for (uint32_t j = 0; j < k; ++j) {
        double acc = 0.0;
        for (uint32_t d = 0; d < D; ++d) {
            const double v = (double)seed_value(row, dtype, d) - cl[j * D + d];
            acc = acc + v * v; // two roundings, in dimension order
        }
        if (j == 0 || acc < best) {
            best = acc;
            bi = j;
        }
    }
    label[i] = bi;
    atomicAdd(&counts[bi], 1u);
    for (uint32_t d = 0; d < D; ++d)
        atomicAdd((unsigned long long *)&sums[bi * D + d],
                  (unsigned long long)(long long)seed_value(row, dtype, d));
}

// centre <- sum / count (a centre with no rows keeps its place); sums and
// counts return to zero for the next step
__global__ void seed_update_kernel(double *__restrict__ c, long long *__restrict__ sums,
                                   uint32_t *__restrict__ counts, uint32_t k, uint32_t D)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * D)
        return;
    const uint32_t m = counts[t / D];
    if (m)
        c[t] = (double)sums[t] / (double)m;
    sums[t] = 0;
}

__global__ void seed_clear_counts_kernel(uint32_t *__restrict__ counts, uint32_t k)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < k)
        counts[t] = 0;
}

// per centre (one workgroup): the nearest sample row, lowest sample index
// on ties
__global__ __launch_bounds__(256) void seed_nearest_kernel(
    const uint8_t *__restrict__ vecs, uint32_t stride, int dtype, const uint32_t *__restrict__ ids,
    uint32_t S, const double *__restrict__ c, uint32_t D, uint32_t *__restrict__ out)
{
    __shared__ double bd[256];
    __shared__ uint32_t bs[256];
    const uint32_t j = blockIdx.x;
    double best = 0.0;
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        const uint8_t *row = vecs + (uint64_t)ids[i] * stride;
        double acc = 0.0;
        for (uint32_t d = 0; d < D; ++d) {
            const double v = (double)seed_value(row, dtype, d) - c[j * D + d];
            acc = acc + v * v; // two roundings, in dimension order
        }
        if (bi == 0xFFFFFFFFu || acc < best) {
            best = acc;
            bi = i;
        }
    }
    bd[threadIdx.x] = best;
    bs[threadIdx.x] = bi;
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const uint32_t o = threadIdx.x + h;
            const bool take = bs[o] != 0xFFFFFFFFu &&
                              (bs[threadIdx.x] == 0xFFFFFFFFu || bd[o] < bd[threadIdx.x] ||
                               (bd[o] == bd[threadIdx.x] && bs[o] < bs[threadIdx.x]));
            if (take) {
                bd[threadIdx.x] = bd[o];
                bs[threadIdx.x] = bs[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[j] = ids[bs[0]];
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_pick_seeds(const nbpg_index *ix, uint32_t count, uint32_t sample,
                               uint32_t iters, uint64_t seed, uint32_t *h_out_ids,
                               uint32_t *out_count)
{
    NBPG_TRY(ensure_init());
    if (!ix || !h_out_ids || !out_count)
        return set_error(NBPG_EINVAL, "null argument");
    *out_count = 0;
    if (count == 0 || count > SEED_MAXK)
        return set_error(NBPG_EINVAL, "count %u (1..%d)", count, SEED_MAXK);
    const uint32_t D = (ix->dtype == NBPG_U4X2 || ix->dtype == NBPG_I4X2) ? 2 * ix->dims : ix->dims;
    if (D > SEED_MAXD)
        return set_error(NBPG_EUNSUP, "rows of %u values (<= %d)", D, SEED_MAXD);
    const uint32_t S = (uint32_t)std::min<uint64_t>(std::max(sample, count), 1u << 24);
    if ((uint64_t)count * D * 8 > (uint64_t)g_dev.lds_max)
        return set_error(NBPG_EUNSUP, "%u centres of %u values exceed LDS", count, D);
    const uint32_t k = count;
    uint32_t *ids = nullptr, *label = nullptr, *counts = nullptr, *near = nullptr;
    double *c = nullptr;
    long long *sums = nullptr;
    auto cleanup = [&]() {
        hipFree(ids);
        hipFree(label);
        hipFree(counts);
        hipFree(near);
        hipFree(c);
        hipFree(sums);
    };
    const bool ok = hipMalloc(&ids, S * 4ull) == hipSuccess &&
                    hipMalloc(&label, S * 4ull) == hipSuccess &&
                    hipMalloc(&counts, k * 4ull) == hipSuccess &&
                    hipMalloc(&near, k * 4ull) == hipSuccess &&
                    hipMalloc(&c, (uint64_t)k * D * 8) == hipSuccess &&
                    hipMalloc(&sums, (uint64_t)k * D * 8) == hipSuccess;
    if (!ok) {
        cleanup();
        return set_error(NBPG_ENOMEM, "seed workspace");
    }
    const uint32_t kd = k * D, b256 = 256;
    hipLaunchKernelGGL(seed_sample_kernel, dim3((S + 255) / 256), dim3(b256), 0, 0, ix->n, S, seed,
                       ids);
    hipLaunchKernelGGL(seed_init_kernel, dim3((kd + 255) / 256), dim3(b256), 0, 0, ix->vecs,
                       ix->stride, ix->dtype, ids, k, D, c);
    hipLaunchKernelGGL(seed_clear_counts_kernel, dim3((k + 255) / 256), dim3(b256), 0, 0, counts, k);
    hipError_t e = hipMemset(sums, 0, (uint64_t)kd * 8);
    for (uint32_t it = 0; it < iters && e == hipSuccess; ++it) {
        hipLaunchKernelGGL(seed_assign_kernel, dim3((S + 255) / 256), dim3(b256), kd * 8, 0,
                           ix->vecs, ix->stride, ix->dtype, ids, S, c, k, D, label, sums, counts);
        hipLaunchKernelGGL(seed_update_kernel, dim3((kd + 255) / 256), dim3(b256), 0, 0, c, sums,
                           counts, k, D);
        hipLaunchKernelGGL(seed_clear_counts_kernel, dim3((k + 255) / 256), dim3(b256), 0, 0,
                           counts, k);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        hipLaunchKernelGGL(seed_nearest_kernel, dim3(k), dim3(b256), 0, 0, ix->vecs, ix->stride,
                           ix->dtype, ids, S, c, D, near);
    std::vector<uint32_t> h(k);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy(h.data(), near, k * 4ull, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess)
        return set_error(NBPG_EHIP, "seed selection: %s", hipGetErrorString(e));
    std::sort(h.begin(), h.end());
    h.erase(std::unique(h.begin(), h.end()), h.end());
    std::copy(h.begin(), h.end(), h_out_ids);
    *out_count = (uint32_t)h.size();
    return NBPG_OK;
}
