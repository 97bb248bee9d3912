// build.hip -- build_vamana (nsw.h:547-663) with every batch on the device.
//
// Per batch (NSWIndex::batch_insert, nsw.h:110-286):
//   insert_kernel   one wavefront per inserted node: the reference-exact
//                   beam search from the node's start points (k = 0, visited
//                   cap VL), existing neighbours the filter has not seen
//                   (nsw.h:150-161), merge_union with the visited list capped
//                   at VL (nsw.h:162-165, utilities.h:128-165),
//                   sort_candidates + RobustPrune (nsw.h:166-170); lists of
//                   up to 256 rows of up to 128 bytes are handed to
//   build_prune_kernel  RobustPrune on int8 matrix cores (same survivors),
//                   one wavefront per inserted node
//   forward_kernel  each node's row := its pruned list (nsw.h:416-429)
//   edges + sorts   reverse edges (target, dist, source), ordered by target
//                   then (dist, source) with two stable radix sorts, grouped
//                   by run-length encoding (nsw.h:431-444, merge_edges.h:53-81)
//   reverse_kernel  one wavefront per target: first VL sources, existing
//                   neighbours without repeats, exact distances, sort and keep
//                   the R closest (nsw.h:463-511)
// Passes, shuffles and start-point policies follow build_vamana; the
// popularity propagation of the CalPopular pass is an O(N R) sequential host
// pass over the downloaded graph (each node's sum feeds later nodes in id
// order); the isolated-node repair (dfs.h:101-253) runs on the device
// (graph_io.hip).  Random start points (fillMergeSP) come from a counter-based
// splitmix64 stream in place of randomInt's random_device-seeded engine.
#include <float.h>
#include <math.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include "prune_core.h"
#include "prune_mfma.h"
#include "ref_search.h"

namespace nbpg {

enum { POP_FIXED_MIN = 1, POP_FIXED_MAX, POP_CAL, POP_PRUNE };
enum { SP_SPLIT = 0, SP_MERGE, SP_DEFAULT, SP_RAND };

__host__ __device__ inline uint64_t sp_stream(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed + (counter + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct InsertArgs {
    const uint8_t *vecs;
    const uint32_t *norms;
    const uint32_t *nbr;
    const uint32_t *deg;
    uint64_t n, m;
    uint32_t stride, R;
    uint32_t mips; // inner-product metric
    const uint32_t *inserts;
    uint64_t floor, bs;
    int sp_mode;
    uint32_t num_split;
    uint64_t sp_seed, sp_ctr;
    uint32_t L, VL, H;
    uint64_t limit;
    const float *alphas, *coss; // per-node schedules (PopularPrune), or NULL
    float alpha_f, cos_f;
    float *popular;
    int cal_popular;
    uint8_t *gscratch;
    uint64_t wave_bytes;
    uint32_t *p_ids;
    float *p_d;
    uint32_t *p_cnt;
    // matrix-core prune hand-off (or NULL): per insert, PM_MAXC sorted keys
    // and their count, NBPG_SENT where the insert pruned in place
    uint64_t *pm_keys;
    uint32_t *pm_cnt;
};

__host__ __device__ inline uint64_t insert_wave_bytes(uint32_t L, uint32_t R, uint32_t VL)
{
    const RefLayout lo = ref_layout(L, R, VL, ref_hash_size(L));
    uint32_t c = 1;
    while (c < VL)
        c <<= 1;
    return lo.total + 8ull * R + 8ull * c + 16;
}

// A wavefront's search state (19 KB at L 70, R 64; 27 KB at L 96) lives in
// LDS when LDS holds ten or more wavefronts per CU or the whole batch at once
// (one wavefront per insert), and otherwise in a global scratch slot per
// wavefront of a persistent grid (ref_global_waves_per_simd): LDS would hold
// 5-8 wavefronts per CU, the slots hold 32 (1M x 128, L 70: 1.65 -> 1.50 s
// per build; 2M x 96, L 96: 6.1 -> 3.8 s).
template <int DT, int NCH, bool LDS>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(LDS ? 1 : ref_global_waves_per_simd(NCH))))
void insert_kernel(InsertArgs a)
{
    extern __shared__ __align__(16) uint8_t lds_ins[];
    uint8_t *st = LDS ? lds_ins : a.gscratch + (uint64_t)blockIdx.x * a.wave_bytes;
    const uint32_t lane = threadIdx.x;
    RefState s = ref_state(st, a.L, a.R, a.VL, a.H);
    const RefLayout lo = ref_layout(a.L, a.R, a.VL, a.H);
    uint64_t *ex = reinterpret_cast<uint64_t *>(st + lo.total);
    uint64_t *cp = ex + a.R;
    const GraphView g = {a.vecs, a.norms, a.nbr, a.stride, a.R, a.mips};
    const uint32_t nsp = a.sp_mode == SP_MERGE ? a.num_split : 1;
    for (uint64_t j = blockIdx.x; j < a.bs; j += gridDim.x) {
        const uint32_t index = a.inserts[a.floor + j];
        uint4 qv[NCH];
        const uint4 *qr = reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)index * a.stride);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            qv[c] = qr[c];
        const uint32_t nqn = a.norms[index];
        uint32_t qs[4 * NCH];
        to_uniform<NCH>(qv, qs);
        auto sp_of = [&](uint32_t t) -> uint32_t {
            switch (a.sp_mode) {
            case SP_SPLIT: return index % a.num_split; // nsw.h:567-574
            case SP_DEFAULT: return 0u;                 // nsw.h:610-616
            case SP_RAND: return (uint32_t)(a.m - index - 1); // nsw.h:601-608
            default: {                                  // nsw.h:584-599
                const uint64_t z = sp_stream(a.sp_seed, a.sp_ctr + j * nsp + t);
                uint32_t rid = (uint32_t)(z % a.n);
                rid = rid / a.num_split * a.num_split;
                uint32_t v = rid + t;
                if (v >= a.n)
                    v -= a.num_split;
                return v;
            }
            }
        };
        const RefOut o = ref_search<DT, NCH>(s, g, qs, nqn, sp_of, nsp, 0u, 0.0, a.limit, a.R,
                                             index, lane);
        if (a.cal_popular) // nsw.h:220-230
            for (uint32_t i = lane; i < o.vs; i += WAVE)
                atomicAdd(&a.popular[key_id(s.vis[i])], 1.0f);
        // existing neighbours the filter has not seen (nsw.h:150-161)
        const uint32_t dg = a.deg[index];
        uint32_t fl = 0;
        for (uint32_t j0 = 0; j0 < dg; j0 += WAVE) {
            const uint32_t jj = j0 + lane;
            const bool valid = jj < dg;
            const uint32_t av = valid ? a.nbr[(uint64_t)index * a.R + jj] : NBPG_SENT;
            const bool seen = filter_chunk(s.hash, s.H, av, valid, lane);
            const bool fresh = valid && !seen;
            const uint64_t um = ballot(fresh);
            if (fresh) {
                const uint4 *r = reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)av * a.stride);
                ex[fl + prefix_count(um)] =
                    make_key(dist_from<DT>(a.norms[av], nqn, dot_row<DT, NCH>(r, qs), a.mips), av);
            }
            fl += (uint32_t)__popcll(um);
            wave_sync();
        }
        // merge_union(existing, visited) capped at VL, in the two-pointer
        // order of utilities.h:128-165 (the existing list is unsorted)
        uint32_t nm = 0;
        if (lane == 0) {
            uint32_t i1 = 0, i2 = 0, k = 0;
            while (i1 < fl && i2 < o.vs && k < a.VL) {
                const uint64_t x = ex[i1], y = s.vis[i2];
                if (x < y) {
                    cp[k++] = x;
                    ++i1;
                } else if (y < x) {
                    cp[k++] = y;
                    ++i2;
                } else {
                    cp[k++] = x;
                    ++i1;
                    ++i2;
                }
            }
            while (i1 < fl && k < a.VL)
                cp[k++] = ex[i1++];
            while (i2 < o.vs && k < a.VL)
                cp[k++] = s.vis[i2++];
            nm = k;
        }
        nm = __shfl(nm, 0);
        wave_sync();
        for (uint32_t i = lane; i < nm; i += WAVE)
            if (key_id(cp[i]) == index)
                cp[i] = make_key(FLT_MAX, index);
        wave_sync();
        const uint32_t mm = wave_sort_candidates(cp, nm, index, lane);
        const float af = a.alphas ? a.alphas[index] : a.alpha_f;
        const float cf = a.coss ? a.coss[index] : a.cos_f;
        if (a.pm_keys && mm <= PM_MAXC && ((double)af > 1e-5 || (double)cf > 1e-5)) {
            // the sorted list goes to the matrix-core prune (build_prune_kernel)
            for (uint32_t i = lane; i < mm; i += WAVE)
                a.pm_keys[j * PM_MAXC + i] = cp[i];
            if (lane == 0)
                a.pm_cnt[j] = mm;
        } else {
            uint32_t comps = 0;
            const uint32_t sel = wave_prune<DT, NCH>(cp, mm, index, a.R, (double)af, (double)cf,
                                                     a.vecs, a.stride, a.norms, a.p_ids + j * a.R,
                                                     a.p_d + j * a.R, comps, lane, a.mips);
            if (lane == 0) {
                a.p_cnt[j] = sel;
                if (a.pm_cnt)
                    a.pm_cnt[j] = NBPG_SENT;
            }
        }
        wave_sync();
    }
}

#define BUILD_PRUNE_WAVES 4

// the inserts' lists that insert_kernel handed off (pm_cnt != NBPG_SENT):
// limitedSimilarityPrune on matrix cores (wave_prune_mfma) with the node's
// alpha and cos (nsw.h:170, 297-324); lists arrive sorted and deduplicated
template <int DT, int NCH, bool ANGLE>
__global__ __launch_bounds__(64 * BUILD_PRUNE_WAVES) __attribute__((amdgpu_waves_per_eu(3)))
void build_prune_kernel(InsertArgs a)
{
    __shared__ uint64_t keys_all[BUILD_PRUNE_WAVES][PM_MAXC];
    __shared__ __align__(16) PruneMfmaLds lds_all[BUILD_PRUNE_WAVES];
    const uint32_t lane = lane_id();
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const uint64_t j = (uint64_t)blockIdx.x * BUILD_PRUNE_WAVES + wv;
    if (j >= a.bs)
        return;
    const uint32_t mm = __builtin_amdgcn_readfirstlane(a.pm_cnt[j]);
    if (mm > PM_MAXC) // pruned in place
        return;
    uint64_t *keys = keys_all[wv];
    for (uint32_t i = lane; i < mm; i += WAVE)
        keys[i] = a.pm_keys[j * PM_MAXC + i];
    wave_sync();
    const uint32_t index = a.inserts[a.floor + j];
    const double af = (double)(a.alphas ? a.alphas[index] : a.alpha_f);
    const double cf = (double)(a.coss ? a.coss[index] : a.cos_f);
    const uint32_t sel =
        wave_prune_mfma<DT, NCH, ANGLE>(keys, mm, index, a.R, af, cf, a.vecs, a.stride,
                                        lds_all[wv], a.p_ids + j * a.R, a.p_d + j * a.R, lane,
                                        a.mips);
    if (lane == 0)
        a.p_cnt[j] = sel;
}

__global__ void forward_kernel(const uint32_t *__restrict__ inserts, uint64_t floor, uint64_t bs,
                               const uint32_t *__restrict__ p_ids, const uint32_t *__restrict__ p_cnt,
                               uint32_t R, uint32_t *__restrict__ nbr, uint32_t *__restrict__ deg)
{
    const uint64_t j = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const uint32_t lane = threadIdx.x % WAVE;
    if (j >= bs)
        return;
    const uint32_t src = inserts[floor + j], c = p_cnt[j];
    for (uint32_t i = lane; i < R; i += WAVE)
        nbr[(uint64_t)src * R + i] = i < c ? p_ids[j * R + i] : NBPG_SENT;
    if (lane == 0)
        deg[src] = c;
}

__global__ void edges_kernel(const uint32_t *__restrict__ inserts, uint64_t floor, uint64_t bs,
                             const uint32_t *__restrict__ p_ids, const float *__restrict__ p_d,
                             const uint32_t *__restrict__ p_cnt, uint32_t R,
                             uint32_t *__restrict__ tgt, uint64_t *__restrict__ key2)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= bs * R)
        return;
    const uint64_t j = idx / R;
    const uint32_t i = (uint32_t)(idx % R);
    if (i < p_cnt[j]) {
        tgt[idx] = p_ids[idx];
        key2[idx] = make_key(p_d[idx], inserts[floor + j]);
    } else {
        tgt[idx] = NBPG_SENT;
        key2[idx] = ~0ull;
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

#define REV_CAP 512

struct ReverseArgs {
    const uint8_t *vecs;
    const uint32_t *norms;
    uint32_t *nbr;
    uint32_t *deg;
    uint32_t stride, R, VL, top_degree;
    const uint32_t *uniq, *counts, *offsets, *num_runs;
    const uint64_t *key2;
    uint64_t max_groups;
    uint32_t mips;
};

template <int DT, int NCH>
__global__ __launch_bounds__(256) void reverse_kernel(ReverseArgs a)
{
    __shared__ uint64_t keys_all[4][REV_CAP];
    const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    uint64_t *keys = keys_all[wv];
    const uint64_t gi = (uint64_t)blockIdx.x * 4 + wv;
    if (gi >= a.max_groups || gi >= *a.num_runs)
        return;
    const uint32_t tgt = a.uniq[gi];
    if (tgt == NBPG_SENT)
        return;
    const uint32_t b = a.offsets[gi];
    const uint32_t pre = min(a.counts[gi], a.VL); // merge_edges.h:106-118
    const uint32_t cap = a.VL + a.R;               // merge_edges.h:32, 126
    // group sources in (dist, source) order; keys hold (ordered dist, id)
    for (uint32_t i = lane; i < pre; i += WAVE) {
        const uint64_t k = a.key2[b + i];
        keys[i] = ((uint64_t)ord_f(key_dist(k)) << 32) | key_id(k);
    }
    wave_sync();
    // add_neighbors_without_repeats (nsw.h:358-378) in row order
    uint32_t cs = pre;
    const uint32_t dg = a.deg[tgt];
    for (uint32_t j0 = 0; j0 < dg; j0 += WAVE) {
        const uint32_t jj = j0 + lane;
        const bool valid = jj < dg;
        const uint32_t nb = valid ? a.nbr[(uint64_t)tgt * a.R + jj] : NBPG_SENT;
        bool exists = false;
        for (uint32_t t = 0; t < cs; ++t)
            exists |= key_id(keys[t]) == nb;
        for (int l = 0; l < WAVE; ++l) { // duplicates earlier in this chunk
            const uint32_t o = __builtin_amdgcn_readlane(nb, l);
            exists |= (l < (int)lane) && o == nb;
        }
        const bool add = valid && !exists;
        const uint64_t am = ballot(add);
        const uint32_t pos = cs + prefix_count(am);
        wave_sync();
        if (add && pos < cap)
            keys[pos] = nb; // distance filled below
        cs = min(cap, cs + (uint32_t)__popcll(am));
        wave_sync();
    }
    // eval_dist_of_new_neighbors (nsw.h:329-355)
    uint4 tv[NCH];
    const uint4 *tr = reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)tgt * a.stride);
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        tv[c] = tr[c];
    const uint32_t nt = a.norms[tgt];
    uint32_t ts[4 * NCH];
    to_uniform<NCH>(tv, ts);
    auto exact = [&](uint32_t id) -> float {
        const uint4 *r = reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)id * a.stride);
        return dist_from<DT>(a.norms[id], nt, dot_row<DT, NCH>(r, ts), a.mips);
    };
    const uint32_t end = min(pre + a.top_degree, cs);
    for (uint32_t t = pre + lane; t < end; t += WAVE) {
        const uint32_t id = key_id(keys[t]);
        keys[t] = ((uint64_t)ord_f(exact(id)) << 32) | id;
    }
    wave_sync();
    if (end < cs) {
        const float bb = exact(key_id(keys[cs - 1]));
        if (lane == 0) {
            const float av = unord_f((uint32_t)(keys[end - 1] >> 32));
            const uint32_t lid = key_id(keys[cs - 1]);
            keys[cs - 1] = ((uint64_t)ord_f(bb) << 32) | lid;
            const float delta = (av - bb) / (float)(cs - end);
            for (uint32_t t = end; t < cs - 1; ++t) {
                const float v = av + (float)(t - end + 1) * delta;
                keys[t] = ((uint64_t)ord_f(v) << 32) | key_id(keys[t]);
            }
        }
        wave_sync();
    }
    // sort; above R also collapse equal ids and truncate (nsw.h:499-509)
    uint32_t n2 = pow2_at_least(cs);
    for (uint32_t i = cs + lane; i < n2; i += WAVE)
        keys[i] = ~0ull;
    wave_sync();
    wave_bitonic(keys, n2, lane);
    uint32_t m = cs;
    if (cs > a.R) {
        m = 0;
        for (uint32_t i0 = 0; i0 < cs; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const uint64_t v = i < cs ? keys[i] : 0;
            const bool keep = i < cs && (i == 0 || key_id(keys[i - 1]) != key_id(v));
            const uint64_t km = ballot(keep);
            wave_sync();
            if (keep)
                keys[m + prefix_count(km)] = v;
            m += (uint32_t)__popcll(km);
            wave_sync();
        }
        m = min(m, a.R);
    }
    for (uint32_t i = lane; i < a.R; i += WAVE)
        a.nbr[(uint64_t)tgt * a.R + i] = i < m ? key_id(keys[i]) : NBPG_SENT;
    if (lane == 0)
        a.deg[tgt] = m;
}

// ---------------------------------------------------------------- host

struct BuildCtx {
    nbpg_index *ix;
    const nbpg_build_params *bp;
    hipStream_t s;
    uint64_t N, maxb;
    uint32_t R, L, VL, H, top_degree;
    uint64_t sp_ctr = 0;
    // device buffers
    uint32_t *d_inserts = nullptr, *p_ids = nullptr, *p_cnt = nullptr;
    float *p_d = nullptr;
    uint32_t *tgtA = nullptr, *tgtB = nullptr, *uniq = nullptr, *counts = nullptr,
             *offsets = nullptr, *num_runs = nullptr;
    uint64_t *keyA = nullptr, *keyB = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_bytes = 0;
    float *popular = nullptr, *alphas = nullptr, *coss = nullptr;
    uint8_t *gscratch = nullptr;
    uint64_t *pm_keys = nullptr; // matrix-core prune hand-off (row shapes that allow it)
    uint32_t *pm_cnt = nullptr;
    uint64_t wave_bytes = 0;
    uint32_t ins_grid = 0;
    uint64_t lds_batch = 0; // batches up to this size run with LDS state
    std::vector<float> h_popular;
    float minPop = 0.0f, maxPop = 1.0f;
};

static void build_free(BuildCtx &c)
{
    hipFree(c.d_inserts);
    hipFree(c.p_ids);
    hipFree(c.p_cnt);
    hipFree(c.p_d);
    hipFree(c.tgtA);
    hipFree(c.tgtB);
    hipFree(c.uniq);
    hipFree(c.counts);
    hipFree(c.offsets);
    hipFree(c.num_runs);
    hipFree(c.keyA);
    hipFree(c.keyB);
    hipFree(c.sort_tmp);
    hipFree(c.popular);
    hipFree(c.alphas);
    hipFree(c.coss);
    hipFree(c.gscratch);
    hipFree(c.pm_keys);
    hipFree(c.pm_cnt);
}

static int download_graph(nbpg_index *ix, std::vector<uint32_t> &nbr, std::vector<uint32_t> &deg)
{
    nbr.resize(ix->n * ix->R);
    deg.resize(ix->n);
    HIP_TRY(hipMemcpy(nbr.data(), ix->nbr, nbr.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(deg.data(), ix->deg, deg.size() * 4, hipMemcpyDeviceToHost));
    return NBPG_OK;
}

// progress lines on stderr when NBPG_BUILD_LOG is set (long builds); the
// batch loop then waits for each batch so the lines track the device
static bool build_log_on()
{
    static const bool on = getenv("NBPG_BUILD_LOG") != nullptr;
    return on;
}
static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

static int batch_insert(BuildCtx &c, const std::vector<uint32_t> &inserts, int pop_mode, int sp_mode)
{
    nbpg_index *ix = c.ix;
    const nbpg_build_params *bp = c.bp;
    const uint64_t m = inserts.size();
    HIP_TRY(hipMemcpyAsync(c.d_inserts, inserts.data(), m * 4, hipMemcpyHostToDevice, c.s));
    float alpha_f = 1.0f, cos_f = (float)bp->min_cos_angle; // nsw.h:297-324
    const float *alphas = nullptr, *coss = nullptr;
    double max_cos = 0.0; // the largest per-node cos (PopularPrune)
    if (pop_mode == POP_FIXED_MIN) {
        alpha_f = (float)bp->max_alpha;
        cos_f = (float)bp->max_cos_angle;
    } else if (pop_mode == POP_PRUNE) {
        const float minA = (float)bp->min_alpha, maxA = (float)bp->max_alpha;
        const float minC = (float)bp->min_cos_angle, maxC = (float)bp->max_cos_angle;
        // both thresholds interpolate on the node's popularity, in float
        std::vector<float> al(c.N), co(c.N);
        for (uint64_t i = 0; i < c.N; ++i) {
            const float cof = (c.h_popular[i] - c.maxPop) / (c.minPop - c.maxPop);
            al[i] = minA + cof * (maxA - minA);
            co[i] = minC + cof * (maxC - minC);
            max_cos = std::max(max_cos, (double)co[i]);
        }
        HIP_TRY(hipMemcpyAsync(c.alphas, al.data(), c.N * 4, hipMemcpyHostToDevice, c.s));
        HIP_TRY(hipMemcpyAsync(c.coss, co.data(), c.N * 4, hipMemcpyHostToDevice, c.s));
        HIP_TRY(hipStreamSynchronize(c.s));
        alphas = c.alphas;
        coss = c.coss;
    }
    if (pop_mode == POP_CAL)
        HIP_TRY(hipMemsetAsync(c.popular, 0, c.N * 4, c.s));

    InsertArgs a;
    memset(&a, 0, sizeof(a));
    a.vecs = ix->vecs;
    a.norms = ix->norms;
    a.mips = ix->mips;
    a.nbr = ix->nbr;
    a.deg = ix->deg;
    a.n = c.N;
    a.m = m;
    a.stride = ix->stride;
    a.R = c.R;
    a.inserts = c.d_inserts;
    a.sp_mode = sp_mode;
    a.num_split = bp->num_split;
    a.sp_seed = bp->sp_seed;
    a.L = c.L;
    a.VL = c.VL;
    a.H = c.H;
    a.limit = bp->limit ? bp->limit : c.N;
    a.alphas = alphas;
    a.coss = coss;
    a.alpha_f = alpha_f;
    a.cos_f = cos_f;
    a.popular = c.popular;
    a.cal_popular = pop_mode == POP_CAL;
    a.gscratch = c.gscratch;
    a.wave_bytes = c.wave_bytes;
    a.p_ids = c.p_ids;
    a.p_d = c.p_d;
    a.p_cnt = c.p_cnt;
    a.pm_keys = c.pm_keys;
    a.pm_cnt = c.pm_cnt;
    // the angle test is evaluated when some node's cos exceeds 1e-5 (prune.h:200)
    const bool pm_angle = (coss ? max_cos : (double)cos_f) > 1e-5;
    const uint32_t nsp = sp_mode == SP_MERGE ? bp->num_split : 1;

    ReverseArgs ra;
    ra.vecs = ix->vecs;
    ra.norms = ix->norms;
    ra.mips = ix->mips;
    ra.nbr = ix->nbr;
    ra.deg = ix->deg;
    ra.stride = ix->stride;
    ra.R = c.R;
    ra.VL = c.VL;
    ra.top_degree = c.top_degree;
    ra.uniq = c.uniq;
    ra.counts = c.counts;
    ra.offsets = c.offsets;
    ra.num_runs = c.num_runs;
    ra.key2 = c.keyA;

    uint64_t inc = 0, count = 0;
    const bool logging = build_log_on();
    double t_last = logging ? now_s() : 0.0;
    while (count < m) {
        // next_batch_range, base 2 (nsw.h:30-47)
        uint64_t fl, ce;
        const double p2 = pow(2.0, (double)inc);
        if (p2 <= (double)c.maxb) {
            fl = (uint64_t)p2 - 1;
            ce = std::min<uint64_t>((uint64_t)pow(2.0, (double)(inc + 1)) - 1, m);
            count = ce;
        } else {
            fl = count;
            ce = std::min<uint64_t>(count + c.maxb, m);
            count = count + (ce - fl);
        }
        const uint64_t bs = ce - fl;
        ++inc;
        if (bs == 0)
            continue;
        a.floor = fl;
        a.bs = bs;
        a.sp_ctr = c.sp_ctr;
        if (sp_mode == SP_MERGE)
            c.sp_ctr += bs * nsp;
        int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
            if (bs <= c.lds_batch) {
                hipLaunchKernelGGL((insert_kernel<DT.value, NCH.value, true>), dim3((uint32_t)bs),
                                   dim3(64), (uint32_t)c.wave_bytes, c.s, a);
            } else {
                const uint32_t grid = (uint32_t)std::min<uint64_t>(bs, c.ins_grid);
                hipLaunchKernelGGL((insert_kernel<DT.value, NCH.value, false>), dim3(grid),
                                   dim3(64), 0, c.s, a);
            }
            if constexpr (prune_mfma_shape<DT.value, NCH.value>::ok) {
                if (a.pm_keys) {
                    const dim3 pg((uint32_t)((bs + BUILD_PRUNE_WAVES - 1) / BUILD_PRUNE_WAVES));
                    if (pm_angle)
                        hipLaunchKernelGGL((build_prune_kernel<DT.value, NCH.value, true>), pg,
                                           dim3(64 * BUILD_PRUNE_WAVES), 0, c.s, a);
                    else
                        hipLaunchKernelGGL((build_prune_kernel<DT.value, NCH.value, false>), pg,
                                           dim3(64 * BUILD_PRUNE_WAVES), 0, c.s, a);
                }
            }
        });
        NBPG_TRY(rc);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(forward_kernel, dim3((uint32_t)((bs + 3) / 4)), dim3(256), 0, c.s,
                           c.d_inserts, fl, bs, c.p_ids, c.p_cnt, c.R, ix->nbr, ix->deg);
        const uint64_t ne = bs * c.R;
        hipLaunchKernelGGL(edges_kernel, dim3((uint32_t)((ne + 255) / 256)), dim3(256), 0, c.s,
                           c.d_inserts, fl, bs, c.p_ids, c.p_d, c.p_cnt, c.R, c.tgtA, c.keyA);
        // (dist, source) then stable by target
        size_t tb = c.sort_bytes;
        HIP_TRY(rocprim::radix_sort_pairs(c.sort_tmp, tb, c.keyA, c.keyB, c.tgtA, c.tgtB, ne, 0u,
                                          64u, c.s));
        tb = c.sort_bytes;
        HIP_TRY(rocprim::radix_sort_pairs(c.sort_tmp, tb, c.tgtB, c.tgtA, c.keyB, c.keyA, ne, 0u,
                                          32u, c.s));
        hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((ne + 255) / 256)), dim3(256), 0, c.s,
                           c.counts, ne, 0u);
        tb = c.sort_bytes;
        HIP_TRY(rocprim::run_length_encode(c.sort_tmp, tb, c.tgtA, (unsigned int)ne, c.uniq,
                                           c.counts, c.num_runs, c.s));
        tb = c.sort_bytes;
        HIP_TRY(rocprim::exclusive_scan(c.sort_tmp, tb, c.counts, c.offsets, 0u, (size_t)ne,
                                        rocprim::plus<uint32_t>(), c.s));
        ra.max_groups = ne;
        rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
            hipLaunchKernelGGL((reverse_kernel<DT.value, NCH.value>),
                               dim3((uint32_t)((ne + 3) / 4)), dim3(256), 0, c.s, ra);
        });
        NBPG_TRY(rc);
        HIP_TRY(hipGetLastError());
        if (logging && now_s() - t_last > 5.0) {
            HIP_TRY(hipStreamSynchronize(c.s));
            fprintf(stderr, "[nbpg build] inserted %llu / %llu\n", (unsigned long long)count,
                    (unsigned long long)m);
            t_last = now_s();
        }
    }
    HIP_TRY(hipStreamSynchronize(c.s));

    if (pop_mode == POP_CAL) { // nsw.h:255-278
        std::vector<uint32_t> nbr, deg;
        NBPG_TRY(download_graph(ix, nbr, deg));
        std::vector<float> pop(c.N), tmp(c.N, 0.0f);
        HIP_TRY(hipMemcpy(pop.data(), c.popular, c.N * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < c.N; ++i) {
            tmp[i] = pop[i];
            for (uint32_t j = 0; j < deg[i]; ++j)
                tmp[nbr[i * c.R + j]] += tmp[i];
        }
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (uint64_t i = 0; i < c.N; ++i) {
            tmp[i] = (float)log((double)(tmp[i] + 1.0f));
            mn = std::min(mn, tmp[i]);
            mx = std::max(mx, tmp[i]);
        }
        c.h_popular.swap(tmp);
        c.minPop = mn;
        c.maxPop = mx;
    }
    return NBPG_OK;
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_build(nbpg_index *ix, const nbpg_build_params *bp, uint64_t *out_isolated)
{
    if (!ix || !bp)
        return set_error(NBPG_EINVAL, "null argument");
    if (bp->R != ix->R)
        return set_error(NBPG_EINVAL, "build R %u != index R %u", bp->R, ix->R);
    if (bp->L == 0 || bp->VL == 0)
        return set_error(NBPG_EINVAL, "L and VL must be positive");
    if (bp->VL + bp->R > REV_CAP)
        return set_error(NBPG_EUNSUP, "VL + R = %u > %d", bp->VL + bp->R, REV_CAP);
    const uint32_t nsp_max = bp->num_split > 1 ? bp->num_split : 1;
    if (nsp_max > bp->L || (bp->num_split > 1 && bp->num_split > ix->n)) // query_params.h:206-210
        return set_error(NBPG_EINVAL, "start points %u > beam size %u", nsp_max, bp->L);
    NBPG_TRY(ensure_init());
    BuildCtx c;
    c.ix = ix;
    c.bp = bp;
    c.s = 0;
    c.N = ix->n;
    c.R = ix->R;
    c.L = bp->L;
    c.VL = bp->VL;
    c.H = ref_hash_size(bp->L);
    c.top_degree = bp->top_degree == 0 ? ix->R : bp->top_degree;
    uint64_t maxb = (uint64_t)(0.02 * (double)(float)c.N); // nsw.h:288-295
    maxb = std::min<uint64_t>(maxb, 1000000);
    if (maxb == 0)
        maxb = 1;
    c.maxb = maxb;
    c.wave_bytes = insert_wave_bytes(c.L, c.R, c.VL);
    const uint64_t ne = maxb * c.R;
    bool ok = hipMalloc(&c.d_inserts, c.N * 4) == hipSuccess &&
              hipMalloc(&c.p_ids, ne * 4) == hipSuccess && hipMalloc(&c.p_d, ne * 4) == hipSuccess &&
              hipMalloc(&c.p_cnt, maxb * 4) == hipSuccess &&
              hipMalloc(&c.tgtA, ne * 4) == hipSuccess && hipMalloc(&c.tgtB, ne * 4) == hipSuccess &&
              hipMalloc(&c.uniq, ne * 4) == hipSuccess && hipMalloc(&c.counts, ne * 4) == hipSuccess &&
              hipMalloc(&c.offsets, ne * 4) == hipSuccess &&
              hipMalloc(&c.num_runs, 16) == hipSuccess && hipMalloc(&c.keyA, ne * 8) == hipSuccess &&
              hipMalloc(&c.keyB, ne * 8) == hipSuccess &&
              hipMalloc(&c.popular, c.N * 4) == hipSuccess &&
              hipMalloc(&c.alphas, c.N * 4) == hipSuccess &&
              hipMalloc(&c.coss, c.N * 4) == hipSuccess;
    if (ok) {
        uint32_t per_simd = 8;
        dispatch_dt_nch(ix->dtype, ix->nch,
                        [&](auto, auto NCH) { per_simd = ref_global_waves_per_simd(NCH.value); });
        c.ins_grid = (uint32_t)std::min<uint64_t>(maxb, (uint64_t)g_dev.cu * 4 * per_simd);
        ok = hipMalloc(&c.gscratch, (uint64_t)c.ins_grid * c.wave_bytes) == hipSuccess;
        // LDS state for batches that are resident at once, and for every
        // batch when LDS holds ten or more wavefronts per CU (the rule of the
        // reference-mode search)
        const uint64_t lds_waves =
            c.wave_bytes <= (uint64_t)std::min(g_dev.lds_max, 65536)
                ? std::min<uint64_t>(REF_LDS_PER_CU / c.wave_bytes, 32)
                : 0;
        c.lds_batch = lds_waves >= 10 ? ~0ull : (uint64_t)g_dev.cu * lds_waves;
    }
    bool pm_shape = false;
    dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        pm_shape = prune_mfma_shape<DT.value, NCH.value>::ok;
    });
    if (pm_shape && ok)
        ok = hipMalloc(&c.pm_keys, maxb * PM_MAXC * 8) == hipSuccess &&
             hipMalloc(&c.pm_cnt, maxb * 4) == hipSuccess;
    if (ok) {
        size_t b1 = 0, b2 = 0, b3 = 0, b4 = 0;
        rocprim::radix_sort_pairs(nullptr, b1, c.keyA, c.keyB, c.tgtA, c.tgtB, ne, 0u, 64u, c.s);
        rocprim::radix_sort_pairs(nullptr, b2, c.tgtB, c.tgtA, c.keyB, c.keyA, ne, 0u, 32u, c.s);
        rocprim::run_length_encode(nullptr, b3, c.tgtA, (unsigned int)ne, c.uniq, c.counts,
                                   c.num_runs, c.s);
        rocprim::exclusive_scan(nullptr, b4, c.counts, c.offsets, 0u, (size_t)ne,
                                rocprim::plus<uint32_t>(), c.s);
        c.sort_bytes = std::max(std::max(b1, b2), std::max(b3, b4));
        ok = hipMalloc(&c.sort_tmp, c.sort_bytes) == hipSuccess;
    }
    if (!ok) {
        build_free(c);
        return set_error(NBPG_ENOMEM, "build workspace");
    }
    std::vector<uint32_t> ins(c.N);
    for (uint64_t i = 0; i < c.N; ++i)
        ins[i] = (uint32_t)i;
    int rc = NBPG_OK;
    int pop = POP_FIXED_MAX;
    if (bp->num_split >= 1) // nsw.h:564-582
        for (int i = 0; i < 2 && rc == NBPG_OK; ++i) {
            std::shuffle(ins.begin(), ins.end(), std::mt19937(i));
            rc = batch_insert(c, ins, pop, SP_SPLIT);
            if (build_log_on())
                fprintf(stderr, "[nbpg build] split pass %d done\n", i);
        }
    for (uint32_t i = 0; i < bp->num_passes && rc == NBPG_OK; ++i) { // nsw.h:618-654
        std::shuffle(ins.begin(), ins.end(), std::mt19937(i));
        if (bp->num_passes >= 2 && i == bp->num_passes - 2 && bp->dynamic_prune)
            pop = POP_CAL;
        else if (i == bp->num_passes - 1)
            pop = (bp->dynamic_prune && bp->num_passes > 1) ? POP_PRUNE : POP_FIXED_MIN;
        int spm;
        if (i != bp->num_passes - 1)
            spm = bp->num_split > 1 ? SP_MERGE : SP_DEFAULT;
        else
            spm = bp->rand_sp ? SP_RAND : SP_DEFAULT;
        rc = batch_insert(c, ins, pop, spm);
        if (build_log_on())
            fprintf(stderr, "[nbpg build] pass %u done\n", i);
    }
    uint64_t iso = 0;
    if (rc == NBPG_OK && bp->repair_isolate) // DFS::repair_isolate_node(G, 0), nsw.h:658-660
        rc = device_repair(ix, 0, &iso);
    if (out_isolated)
        *out_isolated = iso;
    hipDeviceSynchronize();
    build_free(c);
    return rc;
}
