// search.hip -- greedy beam search over the device-resident graph.
//
// NBPG_SEARCH_REF: one 64-lane workgroup per query replaying
//   filtered_beam_search (beamSearch.h:43-171) step for step: the lossy
//   direct-mapped visited filter with overwrite-on-miss in neighbour order
//   (query_params.h:223-230, size query_params.h:243-247), the sorted
//   visited list capped at 8L (beamSearch.h:95-117, 279), the delayed merge
//   (beamSearch.h:119-125), sort+unique+set_union (128-138), the k-th cut
//   (140-149) and the frontier \ visited rescan (154-160).  Result IDs and
//   the {visited, dist_cmps} counters equal the reference's.
// NBPG_SEARCH_FAST: persistent kernel, one wavefront per query pulled from
//   an atomic work counter; LDS-resident lossy visited filter and sorted
//   beam; W beam entries expanded per step; candidates worse than the beam's
//   tail are dropped before the merge.
#include <algorithm>
#include <mutex>

#include "ref_search.h"
#include "search_args.h"

namespace nbpg {

// --------------------------------------------------------------------------
// reference-exact mode: one 64-lane workgroup per query
// --------------------------------------------------------------------------

template <int DT, int NCH, bool LDS>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(LDS ? 1 : ref_global_waves_per_simd(NCH))))
void search_ref_kernel(SearchArgs a)
{
    extern __shared__ __align__(16) uint8_t lds_state[];
    uint8_t *st = LDS ? lds_state : a.gscratch + (uint64_t)blockIdx.x * a.state_bytes;
    RefState s = ref_state(st, a.L, a.R, a.maxv, a.H);
    const GraphView g = {a.vecs, a.norms, a.nbr, a.stride, a.R, a.mips};
    const uint32_t lane = threadIdx.x;
    for (uint64_t qi = blockIdx.x; qi < a.nq; qi += gridDim.x) {
        uint4 qv[NCH];
        load_query<NCH>(a.queries + qi * a.qstride, a.qbytes, qv);
        const uint32_t nqn = __builtin_amdgcn_readfirstlane(norm_regs<DT, NCH>(qv));
        uint32_t qs[4 * NCH];
        to_uniform<NCH>(qv, qs);
        const RefOut o = ref_search<DT, NCH>(
            s, g, qs, nqn, [&](uint32_t t) { return a.sp[t]; }, a.sp_num, a.k, a.cut, a.limit,
            a.dlim, NBPG_SENT, lane);
        const uint64_t *res = o.F;
        uint32_t nres = o.fs;
        uint64_t cmps = o.cmps;
        if (a.exact) {
            const uint8_t *rq = a.raw_q + qi * a.raw_qstride;
            uint64_t *src;
            uint32_t n, kk;
            if (a.exact == 1) {
                // rerank: frontier[0:num_check] re-scored (beamSearch.h:221-236)
                n = (uint32_t)min<uint64_t>((uint64_t)a.k * a.rerank_factor, o.fs);
                for (uint32_t j = lane; j < n; j += WAVE) {
                    const uint32_t id = key_id(o.F[j]);
                    s.cand[j] = make_key(
                        raw_dist(a.raw_dtype, a.raw_base + id * a.raw_row, rq, a.raw_dims, a.mips), id);
                }
                src = s.cand;
                kk = min(a.k, n);
            } else {
                // filtering: the capped visited list carries exact distances
                // (beamSearch.h:106-113), is re-sorted by them (163-168) and
                // cut to min(k, fs, vs) (217, 239-246)
                n = o.vs;
                for (uint32_t j = lane; j < n; j += WAVE) {
                    const uint32_t id = key_id(s.vis[j]);
                    s.vis[j] = make_key(
                        raw_dist(a.raw_dtype, a.raw_base + id * a.raw_row, rq, a.raw_dims, a.mips), id);
                }
                src = s.vis;
                kk = min(a.k, min(o.fs, o.vs));
                cmps += o.nins;
            }
            wave_sync();
            wave_rank_select(src, n, kk, s.cs, lane);
            wave_sync();
            res = s.cs;
            nres = kk;
        }
        // frontier[0:k] (beamSearch.h:249-254), or the whole frontier when
        // out_k = L
        for (uint32_t j = lane; j < a.out_k; j += WAVE) {
            const uint64_t oi = qi * a.out_k + j;
            if (j < nres) {
                a.out_ids[oi] = key_id(res[j]);
                if (a.out_dists)
                    a.out_dists[oi] = key_dist(res[j]);
            } else {
                a.out_ids[oi] = NBPG_SENT;
                if (a.out_dists)
                    a.out_dists[oi] = __int_as_float(0x7f800000);
            }
        }
        if (a.out_stats && lane == 0) {
            a.out_stats[2 * qi] = (uint32_t)o.nvis;
            a.out_stats[2 * qi + 1] = (uint32_t)cmps;
        }
        wave_sync();
    }
}

// --------------------------------------------------------------------------
// fast mode
// --------------------------------------------------------------------------

__device__ __forceinline__ uint32_t count_less_unflag(const uint64_t *a, uint32_t n, uint64_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (unflag(a[mid]) < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__host__ __device__ inline uint32_t fast_wave_bytes(uint32_t Lmax, uint32_t ipl, uint32_t flog2)
{
    // filter | beam x2 | cand x2 | selected ids
    uint32_t b = (4u << flog2) + 16u * Lmax + 2u * 8u * 64u * ipl + 16u;
    return (b + 15u) & ~15u;
}

template <int DT, int NCH, int IPL>
__global__ __launch_bounds__(1024) void search_fast_kernel(SearchArgs a)
{
    extern __shared__ __align__(16) uint8_t lds_fast[];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / WAVE;
    uint8_t *base = lds_fast + (uint64_t)wv * a.wave_bytes;
    const uint32_t FS = 1u << a.flog2;
    uint32_t *filt = reinterpret_cast<uint32_t *>(base);
    uint64_t *beamA = reinterpret_cast<uint64_t *>(base + 4u * FS);
    uint64_t *beamB = beamA + a.Lmax;
    uint64_t *cand = beamB + a.Lmax;
    uint64_t *cs = cand + 64 * IPL;
    uint32_t *sel = reinterpret_cast<uint32_t *>(cs + 64 * IPL);
    const uint32_t R = a.R;

    bool first = true;
    while (true) {
        const uint32_t qi = next_query(a.counter, a.nq, lane, first);
        if (qi >= a.nq)
            break;
        uint32_t L = a.Ls ? a.Ls[qi] : a.L;
        L = max(1u, min(L, a.Lmax));
        uint32_t k = a.ks ? a.ks[qi] : a.k;
        k = min(k, min(L, a.out_k));
        uint4 qv[NCH];
        load_query<NCH>(a.queries + (uint64_t)qi * a.qstride, a.qbytes, qv);
        const uint32_t nqn = __builtin_amdgcn_readfirstlane(norm_regs<DT, NCH>(qv));
        uint32_t qs[4 * NCH];
        to_uniform<NCH>(qv, qs);
        for (uint32_t i = lane; i < FS; i += WAVE)
            filt[i] = NBPG_SENT;
        wave_sync();
        uint64_t *beam = beamA, *beam2 = beamB;
        {
            const uint32_t e = a.entry;
            const uint4 *r = reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)e * a.stride);
            const float d = dist_from<DT>(a.norms[e], nqn, dot_row<DT, NCH>(r, qs), a.mips);
            if (lane == 0) {
                beam[0] = make_key(d, e);
                filt[fast_slot(e, a.flog2)] = e;
            }
        }
        uint32_t bs = 1, nexp = 0, cmps = 1, edges = 0;
        wave_sync();
        while (nexp < a.limit) {
            // pick up to W unexpanded entries, best first
            uint32_t got = 0;
            for (uint32_t i0 = 0; i0 < bs && got < a.W; i0 += WAVE) {
                const uint32_t i = i0 + lane;
                const uint64_t e = i < bs ? beam[i] : FAST_FLAG;
                const bool un = i < bs && !(e & FAST_FLAG);
                const uint64_t m = ballot(un);
                const uint32_t p = prefix_count(m);
                if (un && got + p < a.W) {
                    beam[i] = e | FAST_FLAG;
                    sel[got + p] = key_id(e);
                }
                got = min(a.W, got + (uint32_t)__popcll(m));
            }
            if (got == 0)
                break;
            nexp += got;
            wave_sync();
            const uint64_t worst = bs >= L ? unflag(beam[L - 1]) : ~0ull;
            // gather neighbour ids of the selected nodes and filter
            uint32_t ids[IPL];
            bool fresh[IPL];
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                const uint32_t it = lane + WAVE * t;
                uint32_t av = NBPG_SENT;
                if (it < got * R) {
                    const uint32_t node = sel[it / R];
                    av = a.nbr[(uint64_t)node * R + (it % R)];
                }
                ids[t] = av;
            }
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                bool f = ids[t] != NBPG_SENT;
                edges += (uint32_t)__popcll(ballot(f));
                if (f) {
                    const uint32_t s = fast_slot(ids[t], a.flog2);
                    f = filt[s] != ids[t];
                    if (f)
                        filt[s] = ids[t];
                }
                fresh[t] = f;
            }
            // distances of the fresh neighbours, all loads in flight together
            float dd[IPL];
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                dd[t] = 0.f;
                if (fresh[t]) {
                    const uint4 *r =
                        reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)ids[t] * a.stride);
                    dd[t] = dist_from<DT>(a.norms[ids[t]], nqn, dot_row<DT, NCH>(r, qs), a.mips);
                }
            }
            uint32_t nc = 0;
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                const uint64_t key = make_key(dd[t], ids[t]);
                const bool keep = fresh[t] && key < worst;
                const uint64_t m = ballot(keep);
                cmps += (uint32_t)__popcll(ballot(fresh[t]));
                if (keep)
                    cand[nc + prefix_count(m)] = key;
                nc += (uint32_t)__popcll(m);
            }
            wave_sync();
            if (nc == 0)
                continue;
            // sort, drop duplicates and ids already in the beam
            wave_rank_sort(cand, nc, cs, lane);
            wave_sync();
            uint32_t nu = 0;
            for (uint32_t i0 = 0; i0 < nc; i0 += WAVE) {
                const uint32_t i = i0 + lane;
                const uint64_t v = i < nc ? cs[i] : 0;
                bool keep = i < nc && (i == 0 || cs[i - 1] != v);
                if (keep) {
                    const uint32_t p = count_less_unflag(beam, bs, v);
                    keep = !(p < bs && unflag(beam[p]) == v);
                }
                const uint64_t m = ballot(keep);
                if (keep)
                    cand[nu + prefix_count(m)] = v;
                nu += (uint32_t)__popcll(m);
            }
            wave_sync();
            if (nu == 0)
                continue;
            // merge into the second beam buffer by rank
            for (uint32_t i = lane; i < bs; i += WAVE) {
                const uint64_t e = beam[i];
                const uint32_t p = i + count_less(cand, nu, unflag(e));
                if (p < L)
                    beam2[p] = e;
            }
            for (uint32_t j = lane; j < nu; j += WAVE) {
                const uint64_t v = cand[j];
                const uint32_t p = j + count_less_unflag(beam, bs, v);
                if (p < L)
                    beam2[p] = v;
            }
            bs = min(L, bs + nu);
            {
                uint64_t *t = beam;
                beam = beam2;
                beam2 = t;
            }
            wave_sync();
        }
        for (uint32_t j = lane; j < a.out_k; j += WAVE) {
            const uint64_t o = (uint64_t)qi * a.out_k + j;
            if (j < k && j < bs) {
                const uint64_t e = unflag(beam[j]);
                a.out_ids[o] = key_id(e);
                if (a.out_dists)
                    a.out_dists[o] = key_dist(e);
            } else {
                a.out_ids[o] = NBPG_SENT;
                if (a.out_dists)
                    a.out_dists[o] = __int_as_float(0x7f800000);
            }
        }
        if (a.out_stats && lane == 0) {
            a.out_stats[(uint64_t)a.stats_w * qi] = nexp;
            a.out_stats[(uint64_t)a.stats_w * qi + 1] = cmps;
            if (a.stats_w == 3)
                a.out_stats[3ull * qi + 2] = edges;
        }
        wave_sync();
    }
}


// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------

// full-precision rows of a rerank / filtering search
struct ExactRows {
    int dtype;
    const void *base, *queries;
    uint32_t dims;
    uint64_t qstride;
    uint32_t rerank_factor; // 0: filtering mode
};

// where a reference-mode launch of nq queries keeps its per-query state of
// `state` bytes: LDS when LDS holds about 10 wavefronts per CU or the whole
// launch at once, otherwise global slots at more wavefronts per CU; the
// grid of the launch
static void ref_plan(const nbpg_index *ix, uint64_t state, uint64_t nq, bool *use_lds,
                     uint64_t *grid)
{
    const uint64_t lds_budget = std::min<uint64_t>((uint64_t)g_dev.lds_max, 65536);
    const uint64_t lds_waves =
        state <= lds_budget ? std::min<uint64_t>(32, REF_LDS_PER_CU / state) : 0;
    *use_lds = lds_waves >= 10 || (lds_waves > 0 && nq <= g_dev.cu * lds_waves);
    uint32_t per_simd = 8;
    dispatch_dt_nch(ix->dtype, ix->nch,
                    [&](auto, auto NCH) { per_simd = ref_global_waves_per_simd(NCH.value); });
    *grid = std::min<uint64_t>(nq, (uint64_t)g_dev.cu * (*use_lds ? 32 : 4 * per_simd));
}

// global scratch bytes a reference-mode search of nq queries with these
// parameters needs (0: its state stays in LDS)
static uint64_t ref_scratch_bytes(const nbpg_index *ix, const nbpg_query_params *qp, uint64_t nq)
{
    const uint32_t maxv = (uint32_t)std::min<uint64_t>(8ull * qp->L, 1u << 30);
    const RefLayout lo = ref_layout(qp->L, ix->R, maxv, ref_hash_size(qp->L));
    bool use_lds = true;
    uint64_t grid = 0;
    ref_plan(ix, lo.total, nq, &use_lds, &grid);
    return use_lds ? 0 : grid * lo.total;
}

// fixed: caller-provided scratch (fixed_bytes) used instead of the index's
// workspace when the per-query state does not stay in LDS; the grid shrinks
// to what it holds
static int search_ref(const nbpg_index *ix, const void *d_queries, uint64_t nq, uint64_t qstride,
                      const uint32_t *sp, uint32_t sp_num, const nbpg_query_params *qp,
                      uint32_t out_k, const ExactRows *ex, uint32_t *d_ids, float *d_dists,
                      uint32_t *d_stats, hipStream_t stream, uint8_t *fixed = nullptr,
                      uint64_t fixed_bytes = 0)
{
    SearchArgs a;
    memset(&a, 0, sizeof(a));
    a.vecs = ix->vecs;
    a.norms = ix->norms;
    a.nbr = ix->nbr;
    a.n = ix->n;
    a.stride = ix->stride;
    a.R = ix->R;
    a.queries = (const uint8_t *)d_queries;
    a.qstride = qstride;
    a.nq = nq;
    a.qbytes = ix->dims;
    a.mips = ix->mips;
    for (uint32_t s = 0; s < sp_num; ++s)
        a.sp[s] = sp[s];
    a.sp_num = sp_num;
    a.k = qp->k;
    a.L = qp->L;
    a.dlim = qp->degree_limit;
    a.cut = qp->cut;
    a.limit = qp->limit;
    uint64_t mv = 8ull * qp->L; // beamSearch.h:279, 290
    a.maxv = (uint32_t)std::min<uint64_t>(mv, 1u << 30);
    a.H = ref_hash_size(qp->L);
    a.out_k = out_k;
    a.out_ids = d_ids;
    a.out_dists = d_dists;
    a.out_stats = d_stats;
    if (ex) {
        a.exact = ex->rerank_factor ? 1 : 2;
        a.rerank_factor = ex->rerank_factor;
        a.raw_dtype = (uint32_t)ex->dtype;
        a.raw_dims = ex->dims;
        a.raw_row = (uint64_t)ex->dims * raw_elem_bytes(ex->dtype);
        a.raw_base = (const uint8_t *)ex->base;
        a.raw_q = (const uint8_t *)ex->queries;
        a.raw_qstride = ex->qstride;
    }
    const RefLayout lo = ref_layout(a.L, a.R, a.maxv, a.H);
    a.state_bytes = lo.total;
    bool use_lds = true;
    uint64_t grid = 0;
    ref_plan(ix, lo.total, nq, &use_lds, &grid);
    if (!use_lds && fixed) {
        grid = std::min<uint64_t>(grid, fixed_bytes / lo.total);
        if (grid == 0)
            return set_error(NBPG_EUNSUP, "reference-mode state of %llu bytes per query exceeds "
                             "the %llu-byte scratch", (unsigned long long)lo.total,
                             (unsigned long long)fixed_bytes);
        a.gscratch = fixed;
    } else if (!use_lds) {
        NBPG_TRY(ensure_ws(const_cast<nbpg_index *>(ix), grid * lo.total));
        a.gscratch = (uint8_t *)ix->ws;
    }
    int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        if (use_lds)
            hipLaunchKernelGGL((search_ref_kernel<DT.value, NCH.value, true>), dim3((uint32_t)grid),
                               dim3(64), (uint32_t)lo.total, stream, a);
        else
            hipLaunchKernelGGL((search_ref_kernel<DT.value, NCH.value, false>),
                               dim3((uint32_t)grid), dim3(64), 0, stream, a);
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

// counter: the launch's work counter (a u32 in device memory); launches
// that may run at the same time need their own
int search_fast(const nbpg_index *ix, const void *d_queries, uint64_t nq, uint64_t qstride,
                uint32_t entry, const uint32_t *seeds, uint32_t nseeds, uint32_t L, uint32_t k,
                const uint32_t *d_L, const uint32_t *d_k, uint32_t Lmax, uint32_t out_k,
                uint64_t limit, const nbpg_fast_config *cfg, uint32_t *d_ids, float *d_dists,
                uint32_t *d_stats, uint32_t *counter, hipStream_t stream, uint32_t stats_w,
                uint32_t *vis_out, uint32_t vis_cap, bool vis_cmps, bool qstart)
{
    nbpg_fast_config c = {2, 1, 11, 0, 0};
    if (cfg)
        c = *cfg;
    if (c.waves_per_group < 1 || c.waves_per_group > 16 || c.expand_width < 1 ||
        c.expand_width > 4 || c.filter_log2 < 8 || c.filter_log2 > 14)
        return set_error(NBPG_EINVAL, "fast config out of range");
    const uint32_t items = c.expand_width * ix->R;
    const uint32_t ipl = (items + 63) / 64;
    if (ipl > 4)
        return set_error(NBPG_EUNSUP, "expand_width * R = %u > 256", items);
    SearchArgs a;
    memset(&a, 0, sizeof(a));
    a.vecs = ix->vecs;
    a.norms = ix->norms;
    a.nbr = ix->nbr;
    a.n = ix->n;
    a.stride = ix->stride;
    a.R = ix->R;
    a.queries = (const uint8_t *)d_queries;
    a.qstride = qstride;
    a.nq = nq;
    a.qbytes = ix->dims;
    a.mips = ix->mips;
    a.L = L;
    a.k = k;
    a.Ls = d_L;
    a.ks = d_k;
    a.Lmax = Lmax;
    a.out_k = out_k;
    a.W = c.expand_width;
    a.flog2 = c.filter_log2;
    a.entry = entry;
    a.limit = limit;
    if (nseeds > 1) {
        if (nseeds > 64)
            return set_error(NBPG_EINVAL, "%u seeds (at most 64)", nseeds);
        for (uint32_t i = 0; i < nseeds; ++i) {
            if (seeds[i] >= ix->n)
                return set_error(NBPG_EINVAL, "seed %u >= n", seeds[i]);
            for (uint32_t j = 0; j < i; ++j)
                if (seeds[j] == seeds[i])
                    return set_error(NBPG_EINVAL, "seed %u repeated", seeds[i]);
            a.sp[i] = seeds[i];
        }
        a.nseeds = nseeds;
    }
    // beams of up to 256 entries live in registers (search_reg.hip); larger
    // beams use the LDS-resident kernel below
    const bool reg = Lmax <= 4 * WAVE;
    const uint32_t E = Lmax <= WAVE ? 1 : (Lmax <= 2 * WAVE ? 2 : 4);
    if (reg && ((qstride & 3) || (reinterpret_cast<uintptr_t>(d_queries) & 3)))
        return set_error(NBPG_EINVAL, "query rows must be 4-byte aligned");
    const bool split = c.row_split != 0;
    // register beam: filter (+ the row-split kernel's compact list of 64*ipl ids)
    a.wave_bytes = reg ? (split ? (4u << c.filter_log2) + 256u * ipl
                                : (4u << c.filter_log2) + 16u * ix->nch)
                       : fast_wave_bytes(Lmax, ipl, c.filter_log2);
    a.counter = counter;
    a.out_ids = d_ids;
    a.out_dists = d_dists;
    a.out_stats = d_stats;
    a.stats_w = stats_w;
    a.vis_out = vis_out;
    a.vis_cap = vis_out ? vis_cap : 0;
    a.vis_cmps = vis_out && vis_cmps ? 1u : 0u;
    if (vis_out && (!reg || vis_cap == 0))
        return set_error(NBPG_EUNSUP, "visited lists need L <= 256");
    const uint64_t lds = (uint64_t)a.wave_bytes * c.waves_per_group;
    if (lds > (uint64_t)g_dev.lds_max)
        return set_error(NBPG_EUNSUP, "fast search needs %llu B of LDS per group",
                         (unsigned long long)lds);
    uint32_t groups = c.groups;
    if (groups == 0) {
        // fill the LDS of every CU, at most 24 waves per CU: the gathers are
        // bound by the row-request rate, and beyond 24 waves more queries in
        // flight only lengthen each one (launches are also capped at what the
        // device holds at once)
        uint64_t per_cu = std::max<uint64_t>(1, (uint64_t)g_dev.lds_max / lds);
        per_cu = std::min<uint64_t>(per_cu, std::max<uint32_t>(1, 24 / c.waves_per_group));
        groups = (uint32_t)(per_cu * g_dev.cu);
    }
    uint64_t need = (nq + c.waves_per_group - 1) / c.waves_per_group;
    if (groups > need)
        groups = (uint32_t)std::max<uint64_t>(1, need);
    dim3 g(groups), b(64 * c.waves_per_group);
    if (nseeds > 1 && !reg)
        return set_error(NBPG_EUNSUP, "seeded FAST search needs L <= 256");
    if (qstart) {
        if (!reg || vis_out || !d_ids)
            return set_error(NBPG_EUNSUP, "nearest-seed starts need L <= 256 and the id output");
        a.qstart = 1;
        a.nseeds = 0;
    }
    if (reg)
        return launch_search_reg(ix->dtype, ix->nch, ipl, E, split, g, b, (uint32_t)lds, stream, a);
    int rc = dispatch_dt_nch(ix->dtype, ix->nch, [&](auto DT, auto NCH) {
        switch (ipl) {
        case 1: hipLaunchKernelGGL((search_fast_kernel<DT.value, NCH.value, 1>), g, b, (uint32_t)lds, stream, a); break;
        case 2: hipLaunchKernelGGL((search_fast_kernel<DT.value, NCH.value, 2>), g, b, (uint32_t)lds, stream, a); break;
        case 3: hipLaunchKernelGGL((search_fast_kernel<DT.value, NCH.value, 3>), g, b, (uint32_t)lds, stream, a); break;
        default: hipLaunchKernelGGL((search_fast_kernel<DT.value, NCH.value, 4>), g, b, (uint32_t)lds, stream, a); break;
        }
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

// Scratch of the allocation-free entry point (nbpg_search_rerank_noalloc):
// one array per GPU, loaded with this code object and held until the
// process ends.  Host code never passes `scratch` itself to a kernel (on
// the host the name is only a stand-in); it passes the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device and kept.  Calls on one GPU share the array, so calls that overlap
// in time on one GPU race.
static __device__ __align__(256) uint8_t scratch[NBPG_NOALLOC_SCRATCH_BYTES];

static int scratch_on_current_device(uint8_t **out)
{
    static std::mutex mu;
    static uint8_t *addr[64] = {};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return set_error(NBPG_EUNSUP, "device ordinal %d", dev);
    std::lock_guard<std::mutex> g(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NBPG_EHIP, "scratch address on device %d: %s", dev,
                             hipGetErrorString(e));
        addr[dev] = static_cast<uint8_t *>(p);
    }
    *out = addr[dev];
    return NBPG_OK;
}

} // namespace nbpg

using namespace nbpg;

extern "C" {

static int check_search(const nbpg_index *ix, const void *d_queries, uint64_t nq,
                        uint64_t query_stride, const uint32_t *h_start_points, uint32_t sp_num,
                        const nbpg_query_params *qp, int base_mode, bool want_ids,
                        const uint32_t *d_out_ids)
{
    if (!ix || !qp || (nq && (!d_queries || (want_ids && !d_out_ids))))
        return set_error(NBPG_EINVAL, "null argument");
    if (base_mode != NBPG_SEARCH_REF && base_mode != NBPG_SEARCH_FAST)
        return set_error(NBPG_EINVAL, "search mode %d", base_mode);
    if (qp->L == 0)
        return set_error(NBPG_EINVAL, "beam size L = 0");
    if (qp->k > qp->L) // beamSearch.h:268-273
        return set_error(NBPG_EINVAL, "k = %u > beam size L = %u", qp->k, qp->L);
    if (query_stride < ix->dims)
        return set_error(NBPG_EINVAL, "query stride %llu < row bytes %u",
                         (unsigned long long)query_stride, ix->dims);
    // query_params.h:206-210 for the reference mode; FAST takes up to 64 seeds
    if (sp_num == 0 || !h_start_points ||
        (base_mode == NBPG_SEARCH_REF && (sp_num > 32 || sp_num > qp->L)) || sp_num > 64)
        return set_error(NBPG_EINVAL, "start points: %u (reference mode 1..min(32, L), fast 1..64)",
                         sp_num);
    for (uint32_t s = 0; s < sp_num; ++s)
        if (h_start_points[s] >= ix->n)
            return set_error(NBPG_EINVAL, "start point %u >= n", h_start_points[s]);
    if (base_mode == NBPG_SEARCH_FAST && nq >= 0xFFFFFFFFull)
        return set_error(NBPG_EUNSUP, "nq too large");
    return NBPG_OK;
}

// nbpg_search's work; counter: the FAST launch's work counter (NULL: the
// index's next one); ref_fixed: scratch for reference-mode state beyond LDS
// in place of the index's workspace (NULL: the workspace)
static int search_entry(const nbpg_index *ix, const void *d_queries, uint64_t nq,
                        uint64_t query_stride, const uint32_t *h_start_points, uint32_t sp_num,
                        const nbpg_query_params *qp, int mode, const nbpg_fast_config *fast,
                        uint32_t *d_out_ids, float *d_out_dists, uint32_t *d_out_stats,
                        void *stream, uint32_t *counter, uint8_t *ref_fixed = nullptr,
                        uint64_t ref_fixed_bytes = 0)
{
    const bool frontier = (mode & NBPG_SEARCH_FRONTIER) != 0;
    const bool edges = (mode & NBPG_SEARCH_EDGES) != 0;
    const bool visited = (mode & NBPG_SEARCH_VISITED) != 0;
    const bool nearest = (mode & NBPG_SEARCH_SEED_NEAREST) != 0;
    const int base = mode & ~(NBPG_SEARCH_FRONTIER | NBPG_SEARCH_EDGES | NBPG_SEARCH_VISITED |
                              NBPG_SEARCH_SEED_NEAREST);
    if ((edges || visited || nearest) && base != NBPG_SEARCH_FAST)
        return set_error(NBPG_EINVAL, "NBPG_SEARCH_EDGES / VISITED / SEED_NEAREST are FAST-mode "
                                      "options");
    if (nearest && visited)
        return set_error(NBPG_EINVAL, "NBPG_SEARCH_SEED_NEAREST needs the id output");
    if (visited && frontier)
        return set_error(NBPG_EINVAL, "NBPG_SEARCH_VISITED replaces the frontier output");
    NBPG_TRY(check_search(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, base,
                          frontier || visited || qp->k, d_out_ids));
    if (nq == 0)
        return NBPG_OK;
    if (visited) {
        if (qp->L > 4 * WAVE)
            return set_error(NBPG_EUNSUP, "visited lists need L <= 256");
        return search_fast(ix, d_queries, nq, query_stride, h_start_points[0], h_start_points,
                           sp_num, qp->L, qp->k, nullptr, nullptr, qp->L, qp->k, qp->limit, fast,
                           nullptr, nullptr, d_out_stats, counter ? counter : work_counter(ix),
                           as_stream(stream), edges ? 3 : 2, d_out_ids, NBPG_VISITED_CAP(qp->L));
    }
    const uint32_t out_k = frontier ? qp->L : qp->k;
    if (base == NBPG_SEARCH_REF)
        return search_ref(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, out_k,
                          nullptr, d_out_ids, d_out_dists, d_out_stats, as_stream(stream),
                          ref_fixed, ref_fixed_bytes);
    if (nearest) {
        if (out_k == 0)
            return set_error(NBPG_EINVAL, "NBPG_SEARCH_SEED_NEAREST needs k >= 1");
        if (qp->L > 4 * WAVE)
            return set_error(NBPG_EUNSUP, "nearest-seed starts need L <= 256");
        NBPG_TRY(start_select(ix, d_queries, nq, query_stride, d_out_ids, out_k,
                              as_stream(stream)));
    }
    return search_fast(ix, d_queries, nq, query_stride, h_start_points[0], h_start_points, sp_num,
                       qp->L, out_k, nullptr, nullptr, qp->L, out_k, qp->limit, fast, d_out_ids,
                       d_out_dists, d_out_stats, counter ? counter : work_counter(ix),
                       as_stream(stream), edges ? 3 : 2, nullptr, 0, false, nearest);
}

int nbpg_search(const nbpg_index *ix, const void *d_queries, uint64_t nq, uint64_t query_stride,
                const uint32_t *h_start_points, uint32_t sp_num, const nbpg_query_params *qp,
                int mode, const nbpg_fast_config *fast, uint32_t *d_out_ids, float *d_out_dists,
                uint32_t *d_out_stats, void *stream)
{
    return search_entry(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, mode, fast,
                        d_out_ids, d_out_dists, d_out_stats, stream, nullptr);
}

// nbpg_search_rerank's work; fixed != NULL: every scratch buffer comes from
// the fixed_bytes at `fixed` and nothing is allocated
static int search_rerank_impl(const nbpg_index *ix, const void *d_queries, uint64_t nq,
                              uint64_t query_stride, int raw_dtype, const void *d_raw_base,
                              uint32_t raw_dims, const void *d_raw_queries,
                              uint64_t raw_query_stride, const uint32_t *h_start_points,
                              uint32_t sp_num, const nbpg_query_params *qp,
                              uint32_t rerank_factor, int mode, const nbpg_fast_config *fast,
                              uint32_t *d_out_ids, float *d_out_dists, uint32_t *d_out_stats,
                              void *stream, uint8_t *fixed, uint64_t fixed_bytes)
{
    NBPG_TRY(check_search(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, mode,
                          qp && qp->k, d_out_ids));
    if (raw_dtype < NBPG_U8 || raw_dtype > NBPG_F32)
        return set_error(NBPG_EINVAL, "raw dtype %d", raw_dtype);
    if (raw_dims == 0 || raw_query_stride < (uint64_t)raw_dims * raw_elem_bytes(raw_dtype))
        return set_error(NBPG_EINVAL, "raw dims %u / raw query stride %llu", raw_dims,
                         (unsigned long long)raw_query_stride);
    if (nq && (!d_raw_base || !d_raw_queries))
        return set_error(NBPG_EINVAL, "null raw rows");
    if (mode == NBPG_SEARCH_FAST && rerank_factor == 0 && qp->L > 4 * WAVE)
        return set_error(NBPG_EUNSUP, "FAST filtering mode needs L <= 256");
    if (nq == 0 || qp->k == 0)
        return NBPG_OK;
    hipStream_t s = as_stream(stream);
    if (mode == NBPG_SEARCH_REF) {
        const ExactRows ex = {raw_dtype, d_raw_base, d_raw_queries, raw_dims, raw_query_stride,
                              rerank_factor};
        return search_ref(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, qp->k, &ex,
                          d_out_ids, d_out_dists, d_out_stats, s, fixed, fixed_bytes);
    }
    // FAST: the top L of each beam (rerank) or its visited list (filtering)
    // into scratch, then re-scored
    const bool filtering = rerank_factor == 0;
    const uint32_t width = filtering ? NBPG_VISITED_CAP(qp->L) : qp->L;
    if (!filtering && std::min<uint64_t>((uint64_t)qp->k * rerank_factor, qp->L) > RERANK_MAX)
        return set_error(NBPG_EUNSUP, "min(k * rerank_factor, L) > %d", RERANK_MAX);
    const uint64_t cand_bytes = nq * width * sizeof(uint32_t);
    uint32_t *cand;
    if (fixed) {
        if (cand_bytes > fixed_bytes)
            return set_error(NBPG_EINVAL, "%llu candidates (nq x %u) exceed the %llu-byte scratch",
                             (unsigned long long)(nq * width), width,
                             (unsigned long long)fixed_bytes);
        cand = reinterpret_cast<uint32_t *>(fixed);
    } else {
        NBPG_TRY(ensure_ws(const_cast<nbpg_index *>(ix), cand_bytes));
        cand = (uint32_t *)ix->ws;
    }
    // every query writes all L of its slots (NBPG_SENT past its frontier),
    // so the scratch needs no clearing between calls
    if (filtering) {
        // every visited entry is a candidate: k * factor >= the list width
        NBPG_TRY(search_fast(ix, d_queries, nq, query_stride, h_start_points[0], h_start_points,
                             sp_num, qp->L, qp->k, nullptr, nullptr, qp->L, qp->k, qp->limit,
                             fast, nullptr, nullptr, d_out_stats, work_counter(ix), s, 2, cand,
                             width, true));
        return rerank_launch(raw_dtype, ix->mips, d_raw_base, ix->n, raw_dims, d_raw_queries,
                             raw_query_stride, nq, cand, width, width, qp->k,
                             (width + qp->k - 1) / qp->k, d_out_ids, d_out_dists, s);
    }
    NBPG_TRY(search_fast(ix, d_queries, nq, query_stride, h_start_points[0], h_start_points,
                         sp_num, qp->L, qp->L, nullptr, nullptr, qp->L, qp->L, qp->limit, fast,
                         cand, nullptr, d_out_stats, work_counter(ix), s));
    return rerank_launch(raw_dtype, ix->mips, d_raw_base, ix->n, raw_dims, d_raw_queries, raw_query_stride,
                         nq, cand, qp->L, qp->L, qp->k, rerank_factor, d_out_ids, d_out_dists, s);
}

int nbpg_search_rerank(const nbpg_index *ix, const void *d_queries, uint64_t nq,
                       uint64_t query_stride, int raw_dtype, const void *d_raw_base,
                       uint32_t raw_dims, const void *d_raw_queries, uint64_t raw_query_stride,
                       const uint32_t *h_start_points, uint32_t sp_num,
                       const nbpg_query_params *qp, uint32_t rerank_factor, int mode,
                       const nbpg_fast_config *fast, uint32_t *d_out_ids, float *d_out_dists,
                       uint32_t *d_out_stats, void *stream)
{
    return search_rerank_impl(ix, d_queries, nq, query_stride, raw_dtype, d_raw_base, raw_dims,
                              d_raw_queries, raw_query_stride, h_start_points, sp_num, qp,
                              rerank_factor, mode, fast, d_out_ids, d_out_dists, d_out_stats,
                              stream, nullptr, 0);
}

int nbpg_search_rerank_noalloc(const nbpg_index *ix, const void *d_queries, uint64_t nq,
                               uint64_t query_stride, int raw_dtype, const void *d_raw_base,
                               uint32_t raw_dims, const void *d_raw_queries,
                               uint64_t raw_query_stride, const uint32_t *h_start_points,
                               uint32_t sp_num, const nbpg_query_params *qp,
                               uint32_t rerank_factor, int mode, const nbpg_fast_config *fast,
                               uint32_t *d_out_ids, float *d_out_dists, uint32_t *d_out_stats,
                               void *stream)
{
    uint8_t *ws = nullptr;
    NBPG_TRY(scratch_on_current_device(&ws));
    return search_rerank_impl(ix, d_queries, nq, query_stride, raw_dtype, d_raw_base, raw_dims,
                              d_raw_queries, raw_query_stride, h_start_points, sp_num, qp,
                              rerank_factor, mode, fast, d_out_ids, d_out_dists, d_out_stats,
                              stream, ws, NBPG_NOALLOC_SCRATCH_BYTES);
}

int nbpg_search_uneven(const nbpg_index *ix, const void *d_queries, uint64_t nq,
                       uint64_t query_stride, const uint32_t *d_L, const uint32_t *d_k,
                       uint32_t max_L, uint32_t out_k, const uint32_t *h_start_points,
                       uint32_t sp_num, const nbpg_fast_config *fast, uint32_t *d_out_ids,
                       float *d_out_dists, uint32_t *d_out_stats, void *stream)
{
    if (!ix || !h_start_points || (nq && (!d_queries || !d_out_ids || !d_L || !d_k)))
        return set_error(NBPG_EINVAL, "null argument");
    if (max_L == 0 || out_k == 0 || out_k > max_L)
        return set_error(NBPG_EINVAL, "max_L %u / out_k %u", max_L, out_k);
    if (sp_num == 0 || sp_num > 64)
        return set_error(NBPG_EINVAL, "start points: %u (1..64)", sp_num);
    if (h_start_points[0] >= ix->n)
        return set_error(NBPG_EINVAL, "start point %u >= n", h_start_points[0]);
    if (query_stride < ix->dims)
        return set_error(NBPG_EINVAL, "query stride too small");
    if (nq == 0)
        return NBPG_OK;
    if (nq >= 0xFFFFFFFFull)
        return set_error(NBPG_EUNSUP, "nq too large");
    return search_fast(ix, d_queries, nq, query_stride, h_start_points[0], h_start_points, sp_num,
                       0, 0, d_L, d_k, max_L, out_k, ix->n, fast, d_out_ids, d_out_dists,
                       d_out_stats, work_counter(ix), as_stream(stream));
}

} // extern "C"

// ---------------------------------------------------------------------------
// one search captured as a HIP graph, replayed per batch
// ---------------------------------------------------------------------------

struct nbpg_search_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t *counter = nullptr; // the captured launch's own work counter
    uint8_t *scratch = nullptr;  // reference mode's state beyond LDS, the graph's own
};

namespace nbpg {

// captures launch(stream, counter) into a new graph with its own work
// counter; launch is one search call that has already run eagerly
template <typename F>
static int capture_search(nbpg_search_graph **out, F &&launch, uint8_t *scratch = nullptr)
{
    nbpg_search_graph *g = new nbpg_search_graph();
    g->scratch = scratch; // freed with the graph
    // the graph's own work counter, so replays never share one with other
    // launches on the index
    if (hipMalloc(&g->counter, 256) != hipSuccess || hipMemset(g->counter, 0, 256) != hipSuccess) {
        nbpg_search_graph_destroy(g);
        return set_error(NBPG_ENOMEM, "graph work counter");
    }
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        nbpg_search_graph_destroy(g);
        return set_error(NBPG_EHIP, "capture stream");
    }
    int rc = NBPG_OK;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = set_error(NBPG_EHIP, "graph capture start");
    } else {
        rc = launch(s, g->counter);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc == NBPG_OK && e != hipSuccess)
            rc = set_error(NBPG_EHIP, "graph capture: %s", hipGetErrorString(e));
        g->graph = graph;
        if (rc == NBPG_OK && hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess)
            rc = set_error(NBPG_EHIP, "graph instantiation");
    }
    hipStreamDestroy(s);
    if (rc != NBPG_OK) {
        nbpg_search_graph_destroy(g);
        return rc;
    }
    *out = g;
    return NBPG_OK;
}

} // namespace nbpg

extern "C" {

int nbpg_search_graph_create(nbpg_search_graph **out, const nbpg_index *ix, const void *d_queries,
                             uint64_t nq, uint64_t query_stride, const uint32_t *h_start_points,
                             uint32_t sp_num, const nbpg_query_params *qp, int mode,
                             const nbpg_fast_config *fast, uint32_t *d_out_ids,
                             float *d_out_dists, uint32_t *d_out_stats)
{
    if (!out)
        return set_error(NBPG_EINVAL, "null output handle");
    *out = nullptr;
    NBPG_TRY(ensure_init());
    if (nq == 0)
        return set_error(NBPG_EINVAL, "empty batch");
    // one eager call first: it checks the arguments, so the capture
    // allocates nothing
    NBPG_TRY(nbpg_search(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, mode, fast,
                         d_out_ids, d_out_dists, d_out_stats, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    // reference mode's state beyond LDS goes to scratch the graph owns, not
    // to the index's workspace, which a later call may reallocate
    const int base = mode & ~(NBPG_SEARCH_FRONTIER | NBPG_SEARCH_EDGES | NBPG_SEARCH_VISITED |
                              NBPG_SEARCH_SEED_NEAREST);
    const uint64_t sbytes = base == NBPG_SEARCH_REF ? ref_scratch_bytes(ix, qp, nq) : 0;
    uint8_t *scratch = nullptr;
    if (sbytes && hipMalloc(&scratch, sbytes) != hipSuccess)
        return set_error(NBPG_ENOMEM, "graph scratch of %llu bytes", (unsigned long long)sbytes);
    return capture_search(
        out,
        [&](hipStream_t s, uint32_t *counter) {
            return search_entry(ix, d_queries, nq, query_stride, h_start_points, sp_num, qp, mode,
                                fast, d_out_ids, d_out_dists, d_out_stats, s, counter, scratch,
                                sbytes);
        },
        scratch);
}

int nbpg_search_uneven_graph_create(nbpg_search_graph **out, const nbpg_index *ix,
                                    const void *d_queries, uint64_t nq, uint64_t query_stride,
                                    const uint32_t *d_L, const uint32_t *d_k, uint32_t max_L,
                                    uint32_t out_k, const uint32_t *h_start_points,
                                    uint32_t sp_num, const nbpg_fast_config *fast,
                                    uint32_t *d_out_ids, float *d_out_dists,
                                    uint32_t *d_out_stats)
{
    if (!out)
        return set_error(NBPG_EINVAL, "null output handle");
    *out = nullptr;
    NBPG_TRY(ensure_init());
    if (nq == 0)
        return set_error(NBPG_EINVAL, "empty batch");
    // the eager call checks every argument the capture uses
    NBPG_TRY(nbpg_search_uneven(ix, d_queries, nq, query_stride, d_L, d_k, max_L, out_k,
                                h_start_points, sp_num, fast, d_out_ids, d_out_dists, d_out_stats,
                                nullptr));
    HIP_TRY(hipDeviceSynchronize());
    return capture_search(out, [&](hipStream_t s, uint32_t *counter) {
        return search_fast(ix, d_queries, nq, query_stride, h_start_points[0], h_start_points,
                           sp_num, 0, 0, d_L, d_k, max_L, out_k, ix->n, fast, d_out_ids,
                           d_out_dists, d_out_stats, counter, s);
    });
}

int nbpg_search_graph_launch(nbpg_search_graph *g, void *stream)
{
    if (!g || !g->exec)
        return set_error(NBPG_EINVAL, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, as_stream(stream)));
    return NBPG_OK;
}

int nbpg_search_graph_destroy(nbpg_search_graph *g)
{
    if (!g)
        return NBPG_OK;
    if (g->exec)
        hipGraphExecDestroy(g->exec);
    if (g->graph)
        hipGraphDestroy(g->graph);
    if (g->counter || g->scratch)
        hipDeviceSynchronize(); // a replay still running reads them
    if (g->counter)
        hipFree(g->counter);
    if (g->scratch)
        hipFree(g->scratch);
    delete g;
    return NBPG_OK;
}

} // extern "C"
