// index.hip -- C-ABI: errors, device setup, the device-resident index,
// distance batches.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "internal.h"

namespace nbpg {

static thread_local char t_err[512] = "";
DevInfo g_dev;

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_err, sizeof(t_err), fmt, ap);
    va_end(ap);
    return code;
}

int ensure_init()
{
    if (g_dev.inited)
        return NBPG_OK;
    return nbpg_init(-1);
}

uint32_t nch_for_bytes(uint32_t bytes)
{
    uint32_t c = (bytes + 15) / 16;
    if (c <= 2)
        return 2;
    if (c <= 4)
        return 4;
    if (c <= 6)
        return 6;
    if (c <= 8)
        return 8;
    if (c <= 16)
        return 16;
    return 0;
}

int ensure_ws(nbpg_index *idx, size_t bytes)
{
    if (idx->ws_bytes >= bytes)
        return NBPG_OK;
    if (idx->ws) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(idx->ws));
        idx->ws = nullptr;
        idx->ws_bytes = 0;
    }
    if (hipMalloc(&idx->ws, bytes) != hipSuccess)
        return set_error(NBPG_ENOMEM, "workspace of %zu bytes", bytes);
    idx->ws_bytes = bytes;
    return NBPG_OK;
}

// dst row r (stride bytes, zero padded) = the first `bytes` of src row r
__global__ void spread_rows_kernel(const uint8_t *__restrict__ src, uint64_t src_stride, uint64_t m,
                                   uint32_t bytes, uint8_t *__restrict__ dst, uint32_t stride)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wpr = stride / 4;
    if (t >= m * wpr)
        return;
    const uint64_t r = t / wpr;
    const uint32_t b0 = (uint32_t)(t % wpr) * 4;
    uint32_t v = 0;
    for (uint32_t j = 0; j < 4; ++j)
        if (b0 + j < bytes)
            v |= (uint32_t)src[r * src_stride + b0 + j] << (8 * j);
    reinterpret_cast<uint32_t *>(dst + r * stride)[b0 / 4] = v;
}

// |row|^2 per row in the type's dot units
template <int DT, int NCH>
__global__ void norms_kernel(const uint8_t *__restrict__ vecs, uint32_t stride, uint64_t first,
                             uint64_t count, uint32_t *__restrict__ norms)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const uint4 *r = reinterpret_cast<const uint4 *>(vecs + (first + i) * stride);
    uint4 x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        x[c] = r[c];
    norms[first + i] = norm_regs<DT, NCH>(x);
}

// reference rows [deg, id...] (stride R+1) -> nbr (stride R, SENT padded) + deg
__global__ void graph_in_kernel(const uint32_t *__restrict__ rows, uint64_t first, uint64_t count,
                                uint32_t R, uint64_t n, uint32_t *__restrict__ nbr,
                                uint32_t *__restrict__ deg, uint32_t *__restrict__ bad)
{
    uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    uint32_t l = threadIdx.x % WAVE;
    if (i >= count)
        return;
    const uint32_t *src = rows + i * (uint64_t)(R + 1);
    uint32_t d = src[0];
    if (d > R) {
        if (l == 0)
            atomicOr(bad, 1u);
        d = 0;
    }
    for (uint32_t j = l; j < R; j += WAVE) {
        uint32_t v = j < d ? src[1 + j] : NBPG_SENT;
        if (j < d && v >= n) {
            atomicOr(bad, 2u);
            v = NBPG_SENT;
        }
        nbr[(first + i) * R + j] = v;
    }
    if (l == 0)
        deg[first + i] = d;
}

__global__ void graph_out_kernel(const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ deg,
                                 uint64_t first, uint64_t count, uint32_t R,
                                 uint32_t *__restrict__ rows)
{
    uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    uint32_t l = threadIdx.x % WAVE;
    if (i >= count)
        return;
    uint32_t *dst = rows + i * (uint64_t)(R + 1);
    uint32_t d = deg[first + i];
    if (l == 0)
        dst[0] = d;
    for (uint32_t j = l; j < R; j += WAVE)
        dst[1 + j] = j < d ? nbr[(first + i) * R + j] : 0u;
}

// one lane per id; the query sits in registers of every lane
template <int DT, int NCH>
__global__ void l2_batch_kernel(const uint8_t *__restrict__ vecs, uint32_t stride,
                                const uint32_t *__restrict__ norms, const uint8_t *__restrict__ q,
                                uint32_t qbytes, const uint32_t *__restrict__ ids, uint64_t n,
                                uint64_t nrows, float *__restrict__ out, uint32_t mips)
{
    uint4 qv[NCH];
    const uint32_t *qw = reinterpret_cast<const uint32_t *>(q);
    // the query buffer holds exactly qbytes; pad the registers with zeros
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        uint32_t w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint32_t b0 = (c * 16 + t * 4);
            if (b0 + 4 <= qbytes && (reinterpret_cast<uintptr_t>(q) & 3) == 0)
                w[t] = qw[c * 4 + t];
            else {
                uint32_t v = 0;
                for (uint32_t b = 0; b < 4; ++b)
                    if (b0 + b < qbytes)
                        v |= (uint32_t)q[b0 + b] << (8 * b);
                w[t] = v;
            }
        }
        qv[c] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    const uint32_t nq = __builtin_amdgcn_readfirstlane(norm_regs<DT, NCH>(qv));
    uint32_t qs[4 * NCH];
    to_uniform<NCH>(qv, qs);
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t id = ids[i];
    if (id >= nrows) {
        out[i] = __int_as_float(0x7fc00000); // NaN marks an invalid id
        return;
    }
    const uint4 *r = reinterpret_cast<const uint4 *>(vecs + (uint64_t)id * stride);
    out[i] = dist_from<DT>(norms[id], nq, dot_row<DT, NCH>(r, qs), mips);
}

// arbitrary packed rows: one thread per pair, plain byte loop
__global__ void l2_pairs_kernel(int dtype, const uint8_t *__restrict__ a, const uint8_t *__restrict__ b,
                                uint64_t n, uint32_t dims, float *__restrict__ out)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint8_t *x = a + i * dims, *y = b + i * dims;
    int32_t s = 0;
    if (dtype == NBPG_U8) {
        for (uint32_t j = 0; j < dims; ++j) {
            int32_t t = (int32_t)x[j] - (int32_t)y[j];
            s += t * t;
        }
    } else if (dtype == NBPG_I8) {
        for (uint32_t j = 0; j < dims; ++j) {
            int32_t t = (int32_t)(int8_t)x[j] - (int32_t)(int8_t)y[j];
            s += t * t;
        }
    } else {
        for (uint32_t j = 0; j < dims; ++j) {
            int32_t th = (int32_t)(x[j] & 0xF0) - (int32_t)(y[j] & 0xF0);
            int32_t tl = (int32_t)(uint8_t)(x[j] << 4) - (int32_t)(uint8_t)(y[j] << 4);
            s += th * th + tl * tl;
        }
    }
    out[i] = (float)s;
}

// the flag lines, then the work counters (256 bytes each)
#define INDEX_COUNTER_BYTES (512 + 256 * NBPG_WORK_COUNTERS)

uint32_t *work_counter(const nbpg_index *ix)
{
    nbpg_index *m = const_cast<nbpg_index *>(ix);
    const uint32_t s = __atomic_fetch_add(&m->launch_seq, 1u, __ATOMIC_RELAXED) % NBPG_WORK_COUNTERS;
    return ix->counter + 128 + 64 * s;
}

} // namespace nbpg

using namespace nbpg;

extern "C" {

const char *nbpg_last_error(void) { return t_err; }

int nbpg_init(int device)
{
    int cnt = 0;
    HIP_TRY(hipGetDeviceCount(&cnt));
    if (cnt <= 0)
        return set_error(NBPG_EHIP, "no HIP device");
    if (device < 0)
        HIP_TRY(hipGetDevice(&device));
    if (device >= cnt)
        return set_error(NBPG_EINVAL, "device %d of %d", device, cnt);
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceGetAttribute(&g_dev.cu, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(hipDeviceGetAttribute(&g_dev.lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HIP_TRY(hipDeviceGetAttribute(&g_dev.l2, hipDeviceAttributeL2CacheSize, device));
    g_dev.device = device;
    g_dev.inited = true;
    return NBPG_OK;
}

int nbpg_device_info(int *cu_count, int *lds_per_group, int *l2_bytes)
{
    NBPG_TRY(ensure_init());
    if (cu_count)
        *cu_count = g_dev.cu;
    if (lds_per_group)
        *lds_per_group = g_dev.lds_max;
    if (l2_bytes)
        *l2_bytes = g_dev.l2;
    return NBPG_OK;
}

int nbpg_index_create(nbpg_index **out, int dtype, uint64_t n, uint32_t dims, uint32_t R)
{
    if (!out)
        return set_error(NBPG_EINVAL, "null output handle");
    *out = nullptr;
    NBPG_TRY(ensure_init());
    if (dtype < NBPG_U8 || dtype > NBPG_I4X2)
        return set_error(NBPG_EINVAL, "dtype %d", dtype);
    if (n == 0 || n >= 0x7FFFFFFFull)
        return set_error(NBPG_EINVAL, "n = %llu (1 .. 2^31-2 supported)", (unsigned long long)n);
    uint32_t nch = nch_for_bytes(dims);
    if (dims == 0 || nch == 0)
        return set_error(NBPG_EUNSUP, "row of %u bytes (1..256 supported)", dims);
    if (R == 0 || R > 256)
        return set_error(NBPG_EUNSUP, "R = %u (1..256 supported)", R);
    nbpg_index *ix = new nbpg_index();
    memset(ix, 0, sizeof(*ix));
    ix->dtype = dtype;
    ix->n = n;
    ix->dims = dims;
    ix->R = R;
    ix->nch = nch;
    ix->stride = 16 * nch;
    bool ok = hipMalloc(&ix->vecs, n * ix->stride) == hipSuccess &&
              hipMalloc(&ix->norms, n * 4) == hipSuccess &&
              hipMalloc(&ix->nbr, n * (uint64_t)R * 4) == hipSuccess &&
              hipMalloc(&ix->deg, n * 4) == hipSuccess && hipMalloc(&ix->counter, INDEX_COUNTER_BYTES) == hipSuccess;
    if (!ok) {
        nbpg_index_destroy(ix);
        return set_error(NBPG_ENOMEM, "index of %llu rows", (unsigned long long)n);
    }
    HIP_TRY(hipMemset(ix->vecs, 0, n * ix->stride));
    HIP_TRY(hipMemset(ix->norms, 0, n * 4));
    HIP_TRY(hipMemset(ix->nbr, 0xFF, n * (uint64_t)R * 4));
    HIP_TRY(hipMemset(ix->deg, 0, n * 4));
    HIP_TRY(hipMemset(ix->counter, 0, INDEX_COUNTER_BYTES));
    *out = ix;
    return NBPG_OK;
}

int nbpg_index_set_metric(nbpg_index *ix, int metric)
{
    if (!ix)
        return set_error(NBPG_EINVAL, "null argument");
    if (metric != NBPG_METRIC_L2 && metric != NBPG_METRIC_MIPS)
        return set_error(NBPG_EINVAL, "metric %d", metric);
    ix->mips = metric == NBPG_METRIC_MIPS ? 1u : 0u;
    return NBPG_OK;
}

int nbpg_index_destroy(nbpg_index *ix)
{
    if (!ix)
        return NBPG_OK;
    hipDeviceSynchronize();
    drop_seeds(ix);
    hipFree(ix->vecs);
    hipFree(ix->norms);
    hipFree(ix->nbr);
    hipFree(ix->deg);
    hipFree(ix->ws);
    hipFree(ix->counter);
    delete ix;
    return NBPG_OK;
}

int nbpg_index_info(const nbpg_index *ix, int *dtype, uint64_t *n, uint32_t *dims, uint32_t *R,
                    uint32_t *row_stride)
{
    if (!ix)
        return set_error(NBPG_EINVAL, "null index");
    if (dtype)
        *dtype = ix->dtype;
    if (n)
        *n = ix->n;
    if (dims)
        *dims = ix->dims;
    if (R)
        *R = ix->R;
    if (row_stride)
        *row_stride = ix->stride;
    return NBPG_OK;
}

int nbpg_index_set_points(nbpg_index *ix, const void *host_rows, uint64_t host_stride,
                          uint64_t first, uint64_t count)
{
    if (!ix || !host_rows)
        return set_error(NBPG_EINVAL, "null argument");
    if (host_stride < ix->dims || first + count > ix->n || first + count < first)
        return set_error(NBPG_EINVAL, "rows [%llu, %llu) of %llu, stride %llu",
                         (unsigned long long)first, (unsigned long long)(first + count),
                         (unsigned long long)ix->n, (unsigned long long)host_stride);
    if (count == 0)
        return NBPG_OK;
    drop_seeds(ix); // their copies would no longer match the rows
    // a flat copy only when a host row has no bytes past `dims`: padding
    // bytes (the reference leaves them uninitialised, point_range.h:51-56)
    // must not reach the device rows, whose tails every kernel sums as zeros
    if (host_stride == ix->stride && ix->dims == ix->stride) {
        HIP_TRY(hipMemcpy(ix->vecs + first * ix->stride, host_rows, count * ix->stride,
                          hipMemcpyHostToDevice));
    } else {
        // contiguous host chunks into a device staging buffer, then a kernel
        // spreads the rows to the padded stride (one strided copy of many
        // short rows is far slower than a flat copy)
        const uint64_t rows_per = std::max<uint64_t>(1, (64ull << 20) / host_stride);
        const uint64_t chunk = std::min<uint64_t>(rows_per, count);
        uint8_t *stage = nullptr;
        if (hipMalloc(&stage, chunk * host_stride) != hipSuccess)
            return set_error(NBPG_ENOMEM, "point staging");
        for (uint64_t s0 = 0; s0 < count; s0 += chunk) {
            const uint64_t m = std::min<uint64_t>(chunk, count - s0);
            const hipError_t e = hipMemcpy(stage, (const uint8_t *)host_rows + s0 * host_stride,
                                           m * host_stride, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                hipFree(stage);
                return set_error(NBPG_EHIP, "point upload: %s", hipGetErrorString(e));
            }
            const uint64_t words = m * (ix->stride / 4);
            hipLaunchKernelGGL(spread_rows_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256),
                               0, 0, stage, host_stride, m, ix->dims, ix->vecs + (first + s0) * ix->stride,
                               ix->stride);
        }
        const hipError_t e = hipDeviceSynchronize();
        hipFree(stage);
        if (e != hipSuccess)
            return set_error(NBPG_EHIP, "point spread: %s", hipGetErrorString(e));
    }
    int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        uint32_t blocks = (uint32_t)((count + 255) / 256);
        hipLaunchKernelGGL((norms_kernel<DT.value, NCH.value>), dim3(blocks), dim3(256), 0, 0,
                           ix->vecs, ix->stride, first, count, ix->norms);
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return NBPG_OK;
}

int nbpg_index_set_graph(nbpg_index *ix, const uint32_t *host_rows, uint64_t first, uint64_t count)
{
    if (!ix || !host_rows)
        return set_error(NBPG_EINVAL, "null argument");
    if (first + count > ix->n || first + count < first)
        return set_error(NBPG_EINVAL, "graph rows out of range");
    const uint64_t chunk = 1 << 20;
    uint32_t *tmp = nullptr;
    size_t tb = (size_t)std::min<uint64_t>(count, chunk) * (ix->R + 1) * 4;
    if (count == 0)
        return NBPG_OK;
    if (hipMalloc(&tmp, tb) != hipSuccess)
        return set_error(NBPG_ENOMEM, "graph staging");
    uint32_t *bad = ix->counter + 32;
    hipMemset(bad, 0, 4);
    int rc = NBPG_OK;
    for (uint64_t s = 0; s < count && rc == NBPG_OK; s += chunk) {
        uint64_t c = std::min<uint64_t>(chunk, count - s);
        if (hipMemcpy(tmp, host_rows + s * (ix->R + 1), c * (ix->R + 1) * 4, hipMemcpyHostToDevice) !=
            hipSuccess) {
            rc = set_error(NBPG_EHIP, "graph upload");
            break;
        }
        hipLaunchKernelGGL(graph_in_kernel, dim3((uint32_t)((c + 3) / 4)), dim3(256), 0, 0, tmp,
                           first + s, c, ix->R, ix->n, ix->nbr, ix->deg, bad);
        if (hipGetLastError() != hipSuccess)
            rc = set_error(NBPG_EHIP, "graph conversion launch");
    }
    uint32_t hbad = 0;
    if (rc == NBPG_OK && hipMemcpy(&hbad, bad, 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(NBPG_EHIP, "graph check readback");
    hipFree(tmp);
    if (rc == NBPG_OK && hbad)
        rc = set_error(NBPG_EINVAL, "graph rows invalid (%s%s)", (hbad & 1) ? "degree > R " : "",
                       (hbad & 2) ? "neighbour id >= n" : "");
    return rc;
}

int nbpg_index_get_graph(const nbpg_index *ix, uint32_t *host_rows, uint64_t first, uint64_t count)
{
    if (!ix || !host_rows)
        return set_error(NBPG_EINVAL, "null argument");
    if (first + count > ix->n || first + count < first)
        return set_error(NBPG_EINVAL, "graph rows out of range");
    if (count == 0)
        return NBPG_OK;
    const uint64_t chunk = 1 << 20;
    uint32_t *tmp = nullptr;
    size_t tb = (size_t)std::min<uint64_t>(count, chunk) * (ix->R + 1) * 4;
    if (hipMalloc(&tmp, tb) != hipSuccess)
        return set_error(NBPG_ENOMEM, "graph staging");
    int rc = NBPG_OK;
    for (uint64_t s = 0; s < count; s += chunk) {
        uint64_t c = std::min<uint64_t>(chunk, count - s);
        hipLaunchKernelGGL(graph_out_kernel, dim3((uint32_t)((c + 3) / 4)), dim3(256), 0, 0, ix->nbr,
                           ix->deg, first + s, c, ix->R, tmp);
        if (hipMemcpy(host_rows + s * (ix->R + 1), tmp, c * (ix->R + 1) * 4, hipMemcpyDeviceToHost) !=
            hipSuccess) {
            rc = set_error(NBPG_EHIP, "graph download");
            break;
        }
    }
    hipFree(tmp);
    return rc;
}

int nbpg_l2_batch(const nbpg_index *ix, const void *d_query, const uint32_t *d_ids, uint64_t n,
                  float *d_out, void *stream)
{
    if (!ix || !d_query || (n && (!d_ids || !d_out)))
        return set_error(NBPG_EINVAL, "null argument");
    if (n == 0)
        return NBPG_OK;
    int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        uint32_t blocks = (uint32_t)((n + 255) / 256);
        hipLaunchKernelGGL((l2_batch_kernel<DT.value, NCH.value>), dim3(blocks), dim3(256), 0,
                           as_stream(stream), ix->vecs, ix->stride, ix->norms,
                           (const uint8_t *)d_query, ix->dims, d_ids, n, ix->n, d_out, ix->mips);
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

int nbpg_l2_pairs(int dtype, const void *d_a, const void *d_b, uint64_t n, uint32_t dims,
                  float *d_out, void *stream)
{
    if ((n && (!d_a || !d_b || !d_out)) || dtype < NBPG_U8 || dtype > NBPG_I4X2)
        return set_error(NBPG_EINVAL, "bad argument");
    if (n == 0)
        return NBPG_OK;
    hipLaunchKernelGGL(l2_pairs_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), dtype, (const uint8_t *)d_a, (const uint8_t *)d_b, n, dims,
                       d_out);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

} // extern "C"
