// search_reg.hip -- fast search with the beam held in vector registers.
//
// One wavefront per query (persistent grid: a wave's first query is its grid
// index, later ones come from an atomic work counter).  The beam of
// L <= 64*E entries lives in registers: lane i holds entries i, 64+i, ... as
// (dist, id) keys in ascending order across lanes, bit 31 of the id word
// marking an expanded entry.  Per step:
//   select   the W best unexpanded entries (one ballot per register row)
//   gather   their adjacency rows (one contiguous R*4-byte load each)
//   filter   a lossy direct-mapped visited filter in LDS (as the reference's,
//            query_params.h:223-230, without its lane-order semantics)
//   score    exact distances of the unseen neighbours; |x|^2 is computed
//            from the gathered row, so a distance is exactly one row gather
//   insert   each neighbour that beats the beam tail: rank = ballot count
//            of smaller keys, entries at or past the rank move up one lane
// until no unexpanded entry is left in the beam.  No LDS is touched by the
// beam, so a step has two dependent global round trips and no LDS chains.
#include "search_args.h"

#include <map>
#include <mutex>
#include <tuple>

namespace nbpg {

// lane i <- lane i-1 (lane 0 <- 0) over the whole wavefront: the GFX9 DPP
// wave_shr:1 move, a VALU op, instead of an LDS-routed ds_bpermute
__device__ __forceinline__ uint64_t shfl_up1(uint64_t v)
{
    constexpr int WAVE_SHR1 = 0x138;
    // bound_ctrl: lane 0, which has no source lane, reads 0
    const uint32_t lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)v, WAVE_SHR1, 0xf, 0xf, true);
    const uint32_t hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), WAVE_SHR1, 0xf, 0xf, true);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Inserts key c into the beam (ascending across lanes, entry i of row e in
// lane i of b[e]).  Bit 31 of an entry's id word marks it expanded; keys
// compare without it.  Entries at positions >= L hold keys pushed out of the
// beam: never smaller than the tail, so they can neither count as smaller
// than a candidate that beats the tail nor be selected.
template <int E>
__device__ __forceinline__ void beam_insert(uint64_t (&b)[E], uint64_t c, uint32_t L, uint32_t lane)
{
    bool dup = false;
    uint32_t pos = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint64_t u = unflag(b[e]);
        dup |= ballot(u == c) != 0;
        pos += (uint32_t)__popcll(ballot(u < c));
    }
    if (dup || pos >= L)
        return;
    uint64_t carry = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        uint64_t up = shfl_up1(b[e]);
        if (E > 1) {
            const uint64_t last = readlane64(b[e], WAVE - 1);
            if (lane == 0)
                up = carry;
            carry = last;
        }
        const uint32_t gp = lane + WAVE * e;
        b[e] = gp > pos ? up : (gp == pos ? c : b[e]);
    }
}

template <int DT, int NCH, int IPL, int E>
__global__ __launch_bounds__(1024) void search_reg_kernel(SearchArgs a)
{
    extern __shared__ __align__(16) uint8_t lds_filter[];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / WAVE;
    const uint32_t FS = 1u << a.flog2;
    // per wave: the visited filter, then the query row (16 * NCH bytes)
    uint32_t *filt = reinterpret_cast<uint32_t *>(lds_filter + (uint64_t)wv * a.wave_bytes);
    uint32_t *qw = filt + FS;
    const uint4 *ql = reinterpret_cast<const uint4 *>(qw);
    const uint32_t R = a.R;
    constexpr uint64_t EMPTY = ~0ull;

    bool first = true;
    while (true) {
        const uint32_t qi = next_query(a.counter, a.nq, lane, first);
        if (qi >= a.nq)
            break;
        uint32_t L = a.Ls ? a.Ls[qi] : a.L;
        L = __builtin_amdgcn_readfirstlane(max(1u, min(L, min(a.Lmax, (uint32_t)(WAVE * E)))));
        uint32_t k = a.ks ? a.ks[qi] : a.k;
        k = __builtin_amdgcn_readfirstlane(min(k, min(L, a.out_k)));
        // the query row into this wave's LDS: one word per lane with a
        // per-lane bound; distances read it back as broadcast chunks
        uint32_t nqn;
        {
            const uint8_t *qp = a.queries + (uint64_t)qi * a.qstride;
            if (lane < 4 * NCH) {
                const uint32_t off = lane * 4;
                uint32_t v = 0;
                if (off + 4 <= a.qbytes)
                    v = *reinterpret_cast<const uint32_t *>(qp + off);
                else
                    for (uint32_t bb = 0; off + bb < a.qbytes && bb < 4; ++bb)
                        v |= (uint32_t)qp[off + bb] << (8 * bb);
                qw[lane] = v;
            }
            wave_sync();
            uint4 qv[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                qv[c] = ql[c];
            nqn = __builtin_amdgcn_readfirstlane(norm_regs<DT, NCH>(qv));
        }
        for (uint32_t i = lane; i < FS; i += WAVE)
            filt[i] = NBPG_SENT;
        uint64_t b[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            b[e] = EMPTY;
        wave_sync(); // filter cleared before the start nodes are marked
        // start nodes (the seeds, or the single entry): one distance per
        // lane, then lane r takes the start node of rank r.  One distance
        // call for both cases keeps the query reads in one place.
        const uint32_t ns = a.nseeds > 1 ? a.nseeds : 1u;
        {
            uint32_t sid = lane == 0 ? (a.qstart ? a.out_ids[(uint64_t)qi * a.out_k] : a.entry)
                                     : NBPG_SENT;
            if (a.nseeds > 1)
                for (uint32_t j = 0; j < a.nseeds; ++j)
                    sid = lane == j ? a.sp[j] : sid;
            uint64_t key = EMPTY;
            if (lane < ns) {
                key = make_key(dist_row_lds<DT, NCH>(
                                   reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)sid * a.stride),
                                   ql, nqn, a.mips),
                               sid);
                filt[fast_slot(sid, a.flog2)] = sid;
            }
            uint32_t rank = 0;
            for (uint32_t j = 0; j < ns; ++j)
                rank += readlane64(key, (int)j) < key ? 1u : 0u;
            // lanes past ns all rank ns; lane `rank` learns who holds that rank
            const uint32_t src = (uint32_t)__builtin_amdgcn_ds_permute((int)(rank * 4), (int)lane);
            const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src * 4), (int)(uint32_t)key);
            const uint32_t hi =
                (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src * 4), (int)(uint32_t)(key >> 32));
            b[0] = (lane < ns && lane < L) ? (((uint64_t)hi << 32) | lo) : EMPTY;
        }
        uint32_t cmps = ns;
        uint32_t nexp = 0, edges = 0;
        wave_sync();
        while (nexp < a.limit) {
            // the W best unexpanded entries
            uint32_t got = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t gp = lane + WAVE * e;
                uint64_t m = ballot(gp < L && b[e] != EMPTY && !(b[e] & FAST_FLAG));
                while (m && got < a.W) {
                    const int l = __builtin_ctzll(m);
                    const uint32_t id = __builtin_amdgcn_readlane((uint32_t)b[e], l);
                    s0 = got == 0 ? id : s0;
                    s1 = got == 1 ? id : s1;
                    s2 = got == 2 ? id : s2;
                    s3 = got == 3 ? id : s3;
                    if ((int)lane == l)
                        b[e] |= FAST_FLAG;
                    ++got;
                    m &= m - 1;
                }
            }
            if (got == 0)
                break;
            if (a.vis_out && lane < got && nexp + lane < a.vis_cap)
                a.vis_out[(uint64_t)qi * a.vis_cap + nexp + lane] =
                    lane == 0 ? s0 : (lane == 1 ? s1 : (lane == 2 ? s2 : s3));
            nexp += got;
            // gather adjacency rows, filter
            uint32_t ids[IPL];
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                const uint32_t it = lane + WAVE * t;
                uint32_t av = NBPG_SENT;
                if (it < got * R) {
                    const uint32_t j = it / R;
                    const uint32_t node = j == 0 ? s0 : (j == 1 ? s1 : (j == 2 ? s2 : s3));
                    av = a.nbr[(uint64_t)node * R + (it - j * R)];
                }
                ids[t] = av;
            }
            bool fresh[IPL];
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                bool f = ids[t] != NBPG_SENT;
                edges += (uint32_t)__popcll(ballot(f));
                if (f) {
                    const uint32_t sl = fast_slot(ids[t], a.flog2);
                    f = filt[sl] != ids[t];
                    if (f)
                        filt[sl] = ids[t];
                }
                fresh[t] = f;
            }
            // exact distances of the unseen neighbours, loads in flight together
            uint64_t key[IPL];
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                key[t] = EMPTY;
                if (fresh[t])
                    key[t] = make_key(
                        dist_row_lds<DT, NCH>(
                            reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)ids[t] * a.stride),
                            ql, nqn, a.mips),
                        ids[t]);
                cmps += (uint32_t)__popcll(ballot(fresh[t]));
            }
            // beam tail
            const uint32_t wl = L - 1;
            uint64_t tail = EMPTY;
#pragma unroll
            for (int e = 0; e < E; ++e)
                if ((uint32_t)e == wl / WAVE)
                    tail = readlane64(b[e], (int)(wl % WAVE));
            const uint64_t worst = unflag(tail);
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                uint64_t m = ballot(key[t] < worst);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    beam_insert<E>(b, readlane64(key[t], l), L, lane);
                    m &= m - 1;
                }
            }
            wave_sync();
        }
        if (a.vis_out) {
            // the rest of the list ends it; when the list is re-scored
            // (filtering mode, beamSearch.h:239-246) that counts as distance
            // computations
            const uint32_t nv = min(nexp, a.vis_cap);
            for (uint32_t i = nv + lane; i < a.vis_cap; i += WAVE)
                a.vis_out[(uint64_t)qi * a.vis_cap + i] = NBPG_SENT;
            if (a.vis_cmps)
                cmps += nv;
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t gp = lane + WAVE * e;
            if (a.out_ids && gp < a.out_k) {
                const uint64_t o = (uint64_t)qi * a.out_k + gp;
                const bool ok = gp < k && b[e] != EMPTY;
                const uint64_t v = unflag(b[e]);
                a.out_ids[o] = ok ? key_id(v) : NBPG_SENT;
                if (a.out_dists)
                    a.out_dists[o] = ok ? key_dist(v) : __int_as_float(0x7f800000);
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

// ---------------------------------------------------------------------------
// row-split variant: the unseen neighbours are compacted into an LDS list and
// scored LPR lanes per row, one 16-byte chunk per lane, so one load
// instruction covers RPI whole rows (8 rows of 128 B) instead of one chunk
// of 64 different rows; the per-lane partials |x_c|^2 - 2<x_c, q_c> are
// summed over the row's lanes with ds_swizzle xor steps.
// ---------------------------------------------------------------------------

template <int M> __device__ __forceinline__ uint32_t swz_xor(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (M << 10));
}

template <int NCH> struct row_split_shape {
    static constexpr int LPR = NCH <= 2 ? 2 : (NCH <= 4 ? 4 : (NCH <= 8 ? 8 : 16));
    static constexpr int RPI = WAVE / LPR;
};

template <int LPR> __device__ __forceinline__ uint32_t row_sum(uint32_t p)
{
    p += swz_xor<1>(p);
    if (LPR >= 4)
        p += swz_xor<2>(p);
    if (LPR >= 8)
        p += swz_xor<4>(p);
    if (LPR >= 16)
        p += swz_xor<8>(p);
    return p;
}

template <int DT, int NCH, int IPL, int E>
__global__ __launch_bounds__(1024) void search_split_kernel(SearchArgs a)
{
    constexpr int LPR = row_split_shape<NCH>::LPR, RPI = row_split_shape<NCH>::RPI;
    constexpr int G = 8; // row groups in flight per lane
    extern __shared__ __align__(16) uint8_t lds_split[];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / WAVE;
    const uint32_t FS = 1u << a.flog2;
    uint32_t *filt = reinterpret_cast<uint32_t *>(lds_split + (uint64_t)wv * a.wave_bytes);
    uint32_t *cl = filt + FS;
    const uint32_t R = a.R;
    const uint32_t c = lane % LPR, rsub = lane / LPR;
    constexpr uint64_t EMPTY = ~0ull;

    bool first = true;
    while (true) {
        const uint32_t qi = next_query(a.counter, a.nq, lane, first);
        if (qi >= a.nq)
            break;
        uint32_t L = a.Ls ? a.Ls[qi] : a.L;
        L = __builtin_amdgcn_readfirstlane(max(1u, min(L, min(a.Lmax, (uint32_t)(WAVE * E)))));
        uint32_t k = a.ks ? a.ks[qi] : a.k;
        k = __builtin_amdgcn_readfirstlane(min(k, min(L, a.out_k)));
        // this lane's 16-byte chunk of the query (zero past the row)
        uint4 qc = make_uint4(0, 0, 0, 0);
        {
            const uint8_t *qp = a.queries + (uint64_t)qi * a.qstride + 16 * c;
            uint32_t w[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t v = 0;
                const uint32_t off = 16 * c + 4 * t;
                if (c < NCH && off + 4 <= a.qbytes)
                    v = *reinterpret_cast<const uint32_t *>(qp + 4 * t);
                else if (c < NCH)
                    for (uint32_t bb = 0; bb < 4; ++bb)
                        if (off + bb < a.qbytes)
                            v |= (uint32_t)qp[4 * t + bb] << (8 * bb);
                w[t] = v;
            }
            qc = make_uint4(w[0], w[1], w[2], w[3]);
        }
        uint32_t nqp = 0;
        nqp = dot_traits<DT>::dot(qc.x, qc.x, nqp);
        nqp = dot_traits<DT>::dot(qc.y, qc.y, nqp);
        nqp = dot_traits<DT>::dot(qc.z, qc.z, nqp);
        nqp = dot_traits<DT>::dot(qc.w, qc.w, nqp);
        const uint32_t nqn = __builtin_amdgcn_readfirstlane(row_sum<LPR>(nqp));
        for (uint32_t i = lane; i < FS; i += WAVE)
            filt[i] = NBPG_SENT;
        uint64_t b[E];
#pragma unroll
        for (int e = 0; e < E; ++e)
            b[e] = EMPTY;

        // scores cl[0:nf) and inserts the ones that beat the beam tail
        auto score_list = [&](uint32_t nf) {
            for (uint32_t base = 0; base < nf; base += RPI * G) {
                uint4 x[G];
                uint32_t id[G];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const uint32_t r = base + g * RPI + rsub;
                    id[g] = r < nf ? cl[r] : 0u;
                    x[g] = make_uint4(0, 0, 0, 0);
                    if (r < nf && c < NCH)
                        x[g] = *reinterpret_cast<const uint4 *>(a.vecs + (uint64_t)id[g] * a.stride +
                                                                16 * c);
                }
                const uint32_t wl = L - 1;
                uint64_t tail = EMPTY;
#pragma unroll
                for (int e = 0; e < E; ++e)
                    if ((uint32_t)e == wl / WAVE)
                        tail = readlane64(b[e], (int)(wl % WAVE));
                const uint64_t worst = unflag(tail);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (base + g * RPI >= nf)
                        break; // no rows left in this group (uniform)
                    const uint32_t r = base + g * RPI + rsub;
                    uint32_t xx = 0, xq = 0;
                    xx = dot_traits<DT>::dot(x[g].x, x[g].x, xx);
                    xx = dot_traits<DT>::dot(x[g].y, x[g].y, xx);
                    xx = dot_traits<DT>::dot(x[g].z, x[g].z, xx);
                    xx = dot_traits<DT>::dot(x[g].w, x[g].w, xx);
                    xq = dot_traits<DT>::dot(x[g].x, qc.x, xq);
                    xq = dot_traits<DT>::dot(x[g].y, qc.y, xq);
                    xq = dot_traits<DT>::dot(x[g].z, qc.z, xq);
                    xq = dot_traits<DT>::dot(x[g].w, qc.w, xq);
                    // inner product: the row's lanes sum <x_c, q_c> alone
                    const uint32_t part = row_sum<LPR>(a.mips ? xq : xx - 2u * xq);
                    const float d = a.mips ? -(float)(int32_t)(part * dot_traits<DT>::scale)
                                           : (float)(int32_t)((nqn + part) * dot_traits<DT>::scale);
                    const uint64_t key = make_key(d, id[g]);
                    uint64_t m = ballot(c == 0 && r < nf && key < worst);
                    while (m) {
                        const int l = __builtin_ctzll(m);
                        beam_insert<E>(b, readlane64(key, l), L, lane);
                        m &= m - 1;
                    }
                }
            }
        };

        // start nodes (the seeds, or the single entry) go through the same
        // scoring as the neighbours; the beam starts empty, so all are kept
        wave_sync();
        const uint32_t ns = a.nseeds > 1 ? a.nseeds : 1u;
        {
            uint32_t sid = lane == 0 ? (a.qstart ? a.out_ids[(uint64_t)qi * a.out_k] : a.entry)
                                     : NBPG_SENT;
            if (a.nseeds > 1)
                for (uint32_t j = 0; j < a.nseeds; ++j)
                    sid = lane == j ? a.sp[j] : sid;
            if (lane < ns) {
                cl[lane] = sid;
                filt[fast_slot(sid, a.flog2)] = sid;
            }
        }
        wave_sync();
        score_list(ns);
        uint32_t nexp = 0, cmps = ns, edges = 0;
        wave_sync();
        while (nexp < a.limit) {
            uint32_t got = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t gp = lane + WAVE * e;
                uint64_t m = ballot(gp < L && b[e] != EMPTY && !(b[e] & FAST_FLAG));
                while (m && got < a.W) {
                    const int l = __builtin_ctzll(m);
                    const uint32_t idn = __builtin_amdgcn_readlane((uint32_t)b[e], l);
                    s0 = got == 0 ? idn : s0;
                    s1 = got == 1 ? idn : s1;
                    s2 = got == 2 ? idn : s2;
                    s3 = got == 3 ? idn : s3;
                    if ((int)lane == l)
                        b[e] |= FAST_FLAG;
                    ++got;
                    m &= m - 1;
                }
            }
            if (got == 0)
                break;
            if (a.vis_out && lane < got && nexp + lane < a.vis_cap)
                a.vis_out[(uint64_t)qi * a.vis_cap + nexp + lane] =
                    lane == 0 ? s0 : (lane == 1 ? s1 : (lane == 2 ? s2 : s3));
            nexp += got;
            uint32_t ids[IPL];
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                const uint32_t it = lane + WAVE * t;
                uint32_t av = NBPG_SENT;
                if (it < got * R) {
                    const uint32_t j = it / R;
                    const uint32_t node = j == 0 ? s0 : (j == 1 ? s1 : (j == 2 ? s2 : s3));
                    av = a.nbr[(uint64_t)node * R + (it - j * R)];
                }
                ids[t] = av;
            }
            uint32_t nf = 0;
#pragma unroll
            for (int t = 0; t < IPL; ++t) {
                bool f = ids[t] != NBPG_SENT;
                edges += (uint32_t)__popcll(ballot(f));
                if (f) {
                    const uint32_t sl = fast_slot(ids[t], a.flog2);
                    f = filt[sl] != ids[t];
                    if (f)
                        filt[sl] = ids[t];
                }
                const uint64_t m = ballot(f);
                if (f)
                    cl[nf + prefix_count(m)] = ids[t];
                nf += (uint32_t)__popcll(m);
            }
            cmps += nf;
            wave_sync();
            score_list(nf);
            wave_sync();
        }
        if (a.vis_out) {
            // the rest of the list ends it; when the list is re-scored
            // (filtering mode, beamSearch.h:239-246) that counts as distance
            // computations
            const uint32_t nv = min(nexp, a.vis_cap);
            for (uint32_t i = nv + lane; i < a.vis_cap; i += WAVE)
                a.vis_out[(uint64_t)qi * a.vis_cap + i] = NBPG_SENT;
            if (a.vis_cmps)
                cmps += nv;
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t gp = lane + WAVE * e;
            if (a.out_ids && gp < a.out_k) {
                const uint64_t o = (uint64_t)qi * a.out_k + gp;
                const bool ok = gp < k && b[e] != EMPTY;
                const uint64_t v = unflag(b[e]);
                a.out_ids[o] = ok ? key_id(v) : NBPG_SENT;
                if (a.out_dists)
                    a.out_dists[o] = ok ? key_dist(v) : __int_as_float(0x7f800000);
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

// Launches a persistent kernel with at most as many workgroups as the device
// holds at once (registers and LDS both counted): a wave's first query is its
// grid index, so a workgroup that has to wait for a free slot would start
// its queries late.
// The occupancy query is answered once per (device, kernel, block, LDS) and
// kept: serving loops launch small batches at a high rate.
template <typename K>
static void launch_resident(K kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t stream,
                            const SearchArgs &a)
{
    static std::mutex mu;
    static std::map<std::tuple<int, const void *, uint32_t, uint32_t>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple(dev, reinterpret_cast<const void *>(kernel), block.x, lds);
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            per_cu = it->second;
        } else {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block.x, lds) !=
                hipSuccess)
                per_cu = 0;
            cache[key] = per_cu;
        }
    }
    if (per_cu > 0)
        grid.x = std::min<uint32_t>(grid.x, (uint32_t)per_cu * g_dev.cu);
    hipLaunchKernelGGL(kernel, grid, block, lds, stream, a);
}

int launch_search_reg(int dtype, uint32_t nch, uint32_t ipl, uint32_t E, bool split, dim3 grid,
                      dim3 block, uint32_t lds, hipStream_t stream, const SearchArgs &a)
{
    int rc = dispatch_dt_nch(dtype, nch, [&](auto DT, auto NCH) {
        constexpr int D = DT.value, C = NCH.value;
#define NBPG_REG_LAUNCH(I, EE)                                                                     \
    do {                                                                                           \
        if (split)                                                                                 \
            launch_resident(search_split_kernel<D, C, I, EE>, grid, block, lds, stream, a);        \
        else                                                                                       \
            launch_resident(search_reg_kernel<D, C, I, EE>, grid, block, lds, stream, a);          \
    } while (0)
        const uint32_t sel = (ipl <= 1 ? 0 : ipl <= 2 ? 1 : 2) * 3 + (E <= 1 ? 0 : E <= 2 ? 1 : 2);
        switch (sel) {
        case 0: NBPG_REG_LAUNCH(1, 1); break;
        case 1: NBPG_REG_LAUNCH(1, 2); break;
        case 2: NBPG_REG_LAUNCH(1, 4); break;
        case 3: NBPG_REG_LAUNCH(2, 1); break;
        case 4: NBPG_REG_LAUNCH(2, 2); break;
        case 5: NBPG_REG_LAUNCH(2, 4); break;
        case 6: NBPG_REG_LAUNCH(4, 1); break;
        case 7: NBPG_REG_LAUNCH(4, 2); break;
        default: NBPG_REG_LAUNCH(4, 4); break;
        }
#undef NBPG_REG_LAUNCH
    });
    NBPG_TRY(rc);
    HIP_TRY(hipGetLastError());
    return NBPG_OK;
}

} // namespace nbpg
