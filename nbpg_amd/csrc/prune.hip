// prune.hip -- batched RobustPrune: Prune::sort_candidates (prune.h:49-74)
// followed by Prune::limitedSimilarityPrune (prune.h:158-209), one
// wavefront per node.
//
// Candidates are keyed (dist, id) with the centre node keyed at FLT_MAX, as
// the reference's biased_less does with dir_bias 0; a bitonic sort in LDS
// orders them, adjacent equal ids collapse (dedupe_by_index, prune.h:251-260),
// trailing centre entries are dropped (prune.h:72-73).  The greedy scan then
// takes candidates in order and, per selected p*, computes d(p*, p') for
// every later live candidate on 64 lanes at once, pruning when
// alpha * d(p*, p') <= d(p, p') in double (prune.h:227-230) or, when
// cos_angle > 1e-5, when the angle test fires (prune.h:232-238).
//
// Rows of up to 128 bytes (8-bit) or 64 bytes (packed 4-bit) take the
// matrix-core scan (prune_mfma.h) for nodes with up to 256 candidates; other
// nodes and row sizes take the vector scan below, in a second launch that
// skips the nodes the first one finished.
#include <algorithm>

#include "prune_mfma.h"

namespace nbpg {

#define PRUNE_MAXC 1024
#define PRUNE_WAVES 4

struct PruneArgs {
    const uint8_t *vecs;
    const uint32_t *norms;
    uint64_t n;
    uint32_t stride;
    uint32_t mips; // inner-product metric
    const uint32_t *nodes;
    uint64_t n_nodes;
    const nbpg_iddist *cand;
    const uint64_t *offsets;
    uint32_t R;
    const float *alphas;
    double alpha, cos_angle;
    uint32_t *out_ids;
    float *out_dists;
    uint32_t *out_counts;
    uint32_t *out_comps; // optional distance counts per node
    uint32_t *flag;      // bit 0: a node had > PRUNE_MAXC candidates; bit 1: bad id;
                         // bit 2: a node had > PM_MAXC candidates (matrix-core kernel)
    uint32_t min_n;      // vector kernel: skip nodes with fewer candidates
};

template <int DT, int NCH>
__global__ __launch_bounds__(64 * PRUNE_WAVES) void prune_kernel(PruneArgs a)
{
    __shared__ uint64_t keys_all[PRUNE_WAVES][PRUNE_MAXC];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / WAVE;
    uint64_t *keys = keys_all[wv];
    const uint64_t t = (uint64_t)blockIdx.x * PRUNE_WAVES + wv;
    if (t >= a.n_nodes)
        return;
    const uint32_t p = a.nodes[t];
    const uint64_t b = a.offsets[t], e = a.offsets[t + 1];
    uint32_t *oid = a.out_ids + t * a.R;
    float *od = a.out_dists ? a.out_dists + t * a.R : nullptr;
    if (e < b || e - b > PRUNE_MAXC || p >= a.n) {
        if (lane == 0) {
            atomicOr(a.flag, p >= a.n ? 2u : 1u);
            a.out_counts[t] = 0;
        }
        return;
    }
    const uint32_t n = (uint32_t)(e - b);
    if (n < a.min_n)
        return; // done by the matrix-core kernel
    uint32_t n2 = 1;
    while (n2 < n)
        n2 <<= 1;
    bool bad = false;
    for (uint32_t i = lane; i < n2; i += WAVE) {
        uint64_t k = ~0ull;
        if (i < n) {
            const nbpg_iddist c = a.cand[b + i];
            bad |= c.id >= a.n;
            k = make_key(c.id == p ? FLT_MAX : c.dist, c.id);
        }
        keys[i] = k;
    }
    if (ballot(bad)) {
        if (lane == 0) {
            atomicOr(a.flag, 2u);
            a.out_counts[t] = 0;
        }
        return;
    }
    wave_sync();
    const uint32_t m = wave_sort_candidates(keys, n, p, lane);
    const double alpha = a.alphas ? (double)a.alphas[t] : a.alpha;
    uint32_t comps = 0;
    // up to 256 candidates of up to 128 bytes: their rows in registers
    uint32_t sel;
    if constexpr (NCH <= 8) {
        if (m <= 4 * WAVE)
            sel = wave_prune_regs<DT, NCH, 4>(keys, m, p, a.R, alpha, a.cos_angle, a.vecs, a.stride,
                                              a.norms, oid, od, comps, lane, a.mips);
        else
            sel = wave_prune<DT, NCH>(keys, m, p, a.R, alpha, a.cos_angle, a.vecs, a.stride, a.norms,
                                      oid, od, comps, lane, a.mips);
    } else {
        sel = wave_prune<DT, NCH>(keys, m, p, a.R, alpha, a.cos_angle, a.vecs, a.stride, a.norms, oid,
                                  od, comps, lane, a.mips);
    }
    if (lane == 0) {
        a.out_counts[t] = sel;
        if (a.out_comps)
            a.out_comps[t] = comps;
    }
}

#define PRUNE_MFMA_WAVES 4

// one wavefront per node, up to PM_MAXC candidates: sort_candidates, then the
// matrix-core scan; a node with more candidates sets flag bit 2 and is left
// to the vector kernel
template <int DT, int NCH, bool ANGLE>
__global__ __launch_bounds__(64 * PRUNE_MFMA_WAVES) __attribute__((amdgpu_waves_per_eu(3)))
void prune_mfma_kernel(PruneArgs a)
{
    __shared__ uint64_t keys_all[PRUNE_MFMA_WAVES][PM_MAXC];
    __shared__ __align__(16) PruneMfmaLds lds_all[PRUNE_MFMA_WAVES];
    const uint32_t lane = lane_id();
    // wave-uniform to the compiler, so the scan's branches stay scalar
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    uint64_t *keys = keys_all[wv];
    const uint64_t t = (uint64_t)blockIdx.x * PRUNE_MFMA_WAVES + wv;
    if (t >= a.n_nodes)
        return;
    const uint32_t p = a.nodes[t];
    const uint64_t b = a.offsets[t], e = a.offsets[t + 1];
    uint32_t *oid = a.out_ids + t * a.R;
    float *od = a.out_dists ? a.out_dists + t * a.R : nullptr;
    if (e < b || p >= a.n) {
        if (lane == 0) {
            atomicOr(a.flag, p >= a.n ? 2u : 1u);
            a.out_counts[t] = 0;
        }
        return;
    }
    if (e - b > PM_MAXC) {
        if (lane == 0)
            atomicOr(a.flag, 4u);
        return;
    }
    const uint32_t n = (uint32_t)(e - b);
    uint32_t n2 = 1;
    while (n2 < n)
        n2 <<= 1;
    bool bad = false;
    for (uint32_t i = lane; i < n2; i += WAVE) {
        uint64_t k = ~0ull;
        if (i < n) {
            const nbpg_iddist c = a.cand[b + i];
            bad |= c.id >= a.n;
            k = make_key(c.id == p ? FLT_MAX : c.dist, c.id);
        }
        keys[i] = k;
    }
    if (ballot(bad)) {
        if (lane == 0) {
            atomicOr(a.flag, 2u);
            a.out_counts[t] = 0;
        }
        return;
    }
    wave_sync();
    const uint32_t m = __builtin_amdgcn_readfirstlane(wave_sort_candidates(keys, n, p, lane));
    const double alpha = a.alphas ? (double)a.alphas[t] : a.alpha;
    uint32_t comps = 0, sel;
    if (alpha <= 1e-5 && a.cos_angle <= 1e-5) // prune.h:164-169, no distances
        sel = wave_prune<DT, NCH>(keys, m, p, a.R, alpha, a.cos_angle, a.vecs, a.stride, a.norms,
                                  oid, od, comps, lane, a.mips);
    else
        sel = wave_prune_mfma<DT, NCH, ANGLE>(keys, m, p, a.R, alpha, a.cos_angle, a.vecs,
                                              a.stride, lds_all[wv], oid, od, lane, a.mips);
    if (lane == 0)
        a.out_counts[t] = sel;
}

// the matrix-core kernel where the row size allows (*used); without it the
// vector kernel runs every node, with it only the nodes that set flag bit 2.
// It counts no distance evaluations: callers that ask for them get the
// vector kernel.
int launch_prune_mfma(const nbpg_index *ix, const PruneArgs &a, hipStream_t stream, bool *used)
{
    *used = false;
    if (a.out_comps)
        return NBPG_OK;
    const uint32_t blocks = (uint32_t)((a.n_nodes + PRUNE_MFMA_WAVES - 1) / PRUNE_MFMA_WAVES);
    const bool angle = a.cos_angle > 1e-5;
    int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        if constexpr (prune_mfma_shape<DT.value, NCH.value>::ok) {
            *used = true;
            if (angle)
                hipLaunchKernelGGL((prune_mfma_kernel<DT.value, NCH.value, true>), dim3(blocks),
                                   dim3(64 * PRUNE_MFMA_WAVES), 0, stream, a);
            else
                hipLaunchKernelGGL((prune_mfma_kernel<DT.value, NCH.value, false>), dim3(blocks),
                                   dim3(64 * PRUNE_MFMA_WAVES), 0, stream, a);
        }
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

int launch_prune(const nbpg_index *ix, const PruneArgs &a, hipStream_t stream)
{
    const uint32_t blocks = (uint32_t)((a.n_nodes + PRUNE_WAVES - 1) / PRUNE_WAVES);
    int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        hipLaunchKernelGGL((prune_kernel<DT.value, NCH.value>), dim3(blocks), dim3(64 * PRUNE_WAVES),
                           0, stream, a);
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

} // namespace nbpg

using namespace nbpg;

extern "C" int nbpg_robust_prune_batch(const nbpg_index *ix, const uint32_t *d_nodes,
                                       uint64_t n_nodes, const nbpg_iddist *d_cand,
                                       const uint64_t *d_offsets, uint32_t R,
                                       const float *d_alpha, double alpha, double cos_angle,
                                       uint32_t *d_out_ids, float *d_out_dists,
                                       uint32_t *d_out_counts, void *stream)
{
    if (!ix || (n_nodes && (!d_nodes || !d_offsets || !d_out_ids || !d_out_counts)))
        return set_error(NBPG_EINVAL, "null argument");
    if (R == 0)
        return set_error(NBPG_EINVAL, "R = 0");
    if (n_nodes == 0)
        return NBPG_OK;
    PruneArgs a;
    memset(&a, 0, sizeof(a));
    a.vecs = ix->vecs;
    a.norms = ix->norms;
    a.mips = ix->mips;
    a.n = ix->n;
    a.stride = ix->stride;
    a.nodes = d_nodes;
    a.n_nodes = n_nodes;
    a.cand = d_cand;
    a.offsets = d_offsets;
    a.R = R;
    a.alphas = d_alpha;
    a.alpha = alpha;
    a.cos_angle = cos_angle;
    a.out_ids = d_out_ids;
    a.out_dists = d_out_dists;
    a.out_counts = d_out_counts;
    a.flag = ix->counter + 40;
    hipStream_t s = as_stream(stream);
    HIP_TRY(hipMemsetAsync(a.flag, 0, 4, s));
    bool mf = false;
    NBPG_TRY(launch_prune_mfma(ix, a, s, &mf));
    uint32_t f = 0;
    if (mf) {
        HIP_TRY(hipMemcpyAsync(&f, a.flag, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        a.min_n = PM_MAXC + 1; // the vector kernel only for the nodes left over
    }
    if (!mf || (f & 4u)) {
        NBPG_TRY(launch_prune(ix, a, s));
        HIP_TRY(hipMemcpyAsync(&f, a.flag, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (f & 2)
        return set_error(NBPG_EINVAL, "node or candidate id >= n");
    if (f & 1)
        return set_error(NBPG_EUNSUP, "more than %d candidates for a node", PRUNE_MAXC);
    return NBPG_OK;
}
