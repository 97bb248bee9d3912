// ref_search.h -- the reference-exact beam search of one wavefront.
//
// Replays filtered_beam_search (beamSearch.h:43-171) in the pure-quantized
// mode step for step, shared by the query kernel (search.hip) and the
// insert kernel of the device build (build.hip):
//   * lossy direct-mapped visited filter, overwrite on miss, neighbours in
//     row order (query_params.h:223-230); size 2^max(10, ceil(log2 L^2) - 2)
//     (query_params.h:243-247)
//   * distances for the unseen neighbours (beamSearch.h:85-93)
//   * sorted insertion into a visited list capped at maxv (beamSearch.h:95-117)
//   * delayed merge while limit >= 2L and fewer than L/8 candidates
//     (beamSearch.h:119-125)
//   * sort + unique by id + set_union with the frontier, capped at L
//     (beamSearch.h:128-138), the k-th entry cut (140-149) and the
//     frontier \ visited rescan (154-160)
// All per-query state sits in one contiguous block (LDS, or global memory
// when it does not fit) laid out by ref_layout().
#pragma once
#include "internal.h"

namespace nbpg {

template <int NCH>
__device__ __forceinline__ void load_query(const uint8_t *__restrict__ q, uint32_t qbytes,
                                           uint4 (&qv)[NCH])
{
    const bool aligned = (reinterpret_cast<uintptr_t>(q) & 15) == 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        if (aligned && (uint32_t)(c * 16 + 16) <= qbytes) {
            qv[c] = reinterpret_cast<const uint4 *>(q)[c];
        } else {
            uint32_t w[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t v = 0;
                for (uint32_t b = 0; b < 4; ++b) {
                    uint32_t off = c * 16 + t * 4 + b;
                    if (off < qbytes)
                        v |= (uint32_t)q[off] << (8 * b);
                }
                w[t] = v;
            }
            qv[c] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
}

struct RefLayout {
    uint64_t fa, fb, unv, cand, cs, vis, total;
};

// hash | frontier x2 (2L+R) | unvisited (L) | candidates x2 (L+R+64) | visited (maxv)
__host__ __device__ inline RefLayout ref_layout(uint32_t L, uint32_t R, uint32_t maxv, uint32_t H)
{
    RefLayout o;
    uint64_t p = ((uint64_t)H * 4 + 15) & ~15ull;
    const uint64_t fsz = 2ull * L + R, csz = (uint64_t)L + R + 64;
    o.fa = p;
    p += 8 * fsz;
    o.fb = p;
    p += 8 * fsz;
    o.unv = p;
    p += 8ull * L;
    o.cand = p;
    p += 8 * csz;
    o.cs = p;
    p += 8 * csz;
    o.vis = p;
    p += 8ull * maxv;
    o.total = (p + 15) & ~15ull;
    return o;
}

// Where a wavefront's state lives.  In LDS, 160 KB per CU (gfx950) holds
// 160 KB / state wavefronts; in a global scratch slot per wavefront, a
// persistent grid holds 8 per SIMD (64 VGPRs; 4 for rows over 128 bytes,
// whose query takes 64 more).  The search's steps are dependent row and
// adjacency gathers, so wavefronts in flight count for more than LDS
// latency once LDS holds fewer than about 10 per CU (1M x 128, 10k
// queries, L 128 (34 KB): 12.1 -> 6.5 ms; L 64 (15 KB): equal).
#define REF_LDS_PER_CU (160u * 1024u)
__host__ __device__ constexpr uint32_t ref_global_waves_per_simd(int NCH) { return NCH > 8 ? 4u : 8u; }

// 2^max(10, ceil(log2(L*L)) - 2) with the reference's uint32 product
__host__ __device__ inline uint32_t ref_hash_size(uint32_t L)
{
    const uint32_t sq = L * L;
    int bits = 0;
    while ((1ull << bits) < (uint64_t)sq)
        ++bits;
    bits -= 2;
    if (bits < 10)
        bits = 10;
    return 1u << bits;
}

// dst[rank(src[i])] = src[i]; stable for equal keys; one wave
__device__ __forceinline__ void wave_rank_sort(const uint64_t *src, uint32_t n, uint64_t *dst,
                                               uint32_t lane)
{
    for (uint32_t i0 = 0; i0 < n; i0 += WAVE) {
        const uint32_t i = i0 + lane;
        const uint64_t v = i < n ? src[i] : 0;
        uint32_t r = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const uint64_t u = src[j];
            r += (u < v) || (u == v && j < i);
        }
        if (i < n)
            dst[r] = v;
    }
}

// the kk smallest of n distinct keys, in order: dst[rank(src[i])] = src[i]
// for rank < kk; one wave
__device__ __forceinline__ void wave_rank_select(const uint64_t *src, uint32_t n, uint32_t kk,
                                                 uint64_t *dst, uint32_t lane)
{
    for (uint32_t i0 = 0; i0 < n; i0 += WAVE) {
        const uint32_t i = i0 + lane;
        const uint64_t v = i < n ? src[i] : 0;
        uint32_t r = 0;
        for (uint32_t j = 0; j < n; ++j)
            r += src[j] < v;
        if (i < n && r < kk)
            dst[r] = v;
    }
}

// has_been_seen (query_params.h:223-230) applied to the valid lanes of one
// 64-id chunk as if called lane by lane in order: a lane sees the id the
// nearest lower lane of its slot wrote, or the slot's prior content.
// Returns "seen"; the last lane of each slot group writes the slot.
__device__ __forceinline__ bool filter_chunk(uint32_t *hash, uint32_t H, uint32_t av, bool valid,
                                             uint32_t lane)
{
    const uint64_t vm = ballot(valid);
    const uint32_t slot = valid ? (uint32_t)(hash64_2(av) & (H - 1)) : 0u;
    const uint32_t hold = valid ? hash[slot] : 0u;
    bool has_pred = false, is_last = true;
    uint32_t pred_a = 0;
    for (uint64_t m = vm; m; m &= m - 1) {
        const int jl = __builtin_ctzll(m);
        const uint32_t sj = __builtin_amdgcn_readlane(slot, jl);
        const uint32_t aj = __builtin_amdgcn_readlane(av, jl);
        if (valid && sj == slot) {
            if (jl < (int)lane) {
                has_pred = true;
                pred_a = aj;
            } else if (jl > (int)lane) {
                is_last = false;
            }
        }
    }
    const bool seen = has_pred ? (pred_a == av) : (hold == av);
    wave_sync();
    if (valid && is_last)
        hash[slot] = av;
    wave_sync();
    return seen;
}

struct RefState {
    uint32_t *hash;
    uint64_t *F, *NF, *unv, *cand, *cs, *vis;
    uint32_t L, R, H, maxv;
};

__device__ __forceinline__ RefState ref_state(uint8_t *st, uint32_t L, uint32_t R, uint32_t maxv,
                                              uint32_t H)
{
    const RefLayout lo = ref_layout(L, R, maxv, H);
    RefState s;
    s.hash = reinterpret_cast<uint32_t *>(st);
    s.F = reinterpret_cast<uint64_t *>(st + lo.fa);
    s.NF = reinterpret_cast<uint64_t *>(st + lo.fb);
    s.unv = reinterpret_cast<uint64_t *>(st + lo.unv);
    s.cand = reinterpret_cast<uint64_t *>(st + lo.cand);
    s.cs = reinterpret_cast<uint64_t *>(st + lo.cs);
    s.vis = reinterpret_cast<uint64_t *>(st + lo.vis);
    s.L = L;
    s.R = R;
    s.H = H;
    s.maxv = maxv;
    return s;
}

struct RefOut {
    uint64_t *F; // final frontier (sorted), fs entries
    uint32_t fs, vs;
    uint64_t nvis, cmps;
    uint64_t nins; // pops that landed in the capped visited list
};

struct GraphView {
    const uint8_t *vecs;
    const uint32_t *norms;
    const uint32_t *nbr;
    uint32_t stride, R;
    uint32_t mips; // inner-product metric
};

// One search.  `self` is skipped among neighbours (beamSearch.h:79, the
// build's same_as test); pass 0xFFFFFFFF for external queries.  `sp_of(s)`
// yields the s-th start point.
template <int DT, int NCH, typename SpFn>
__device__ RefOut ref_search(RefState &s, const GraphView &g, const uint32_t (&qv)[4 * NCH], uint32_t nqn,
                             SpFn sp_of, uint32_t sp_num, uint32_t k, double cut, uint64_t limit,
                             uint32_t dlim, uint32_t self, uint32_t lane)
{
    const uint32_t L = s.L, R = g.R, H = s.H, maxv = s.maxv;
    const uint32_t lim = min(R, dlim);
    auto dist_of = [&](uint32_t id) -> float {
        const uint4 *r = reinterpret_cast<const uint4 *>(g.vecs + (uint64_t)id * g.stride);
        return dist_from<DT>(g.norms[id], nqn, dot_row<DT, NCH>(r, qv), g.mips);
    };
    uint64_t *Fp = s.F, *NFp = s.NF;
    for (uint32_t i = lane; i < H; i += WAVE)
        s.hash[i] = NBPG_SENT;
    wave_sync();
    // init_starting_points (query_params.h:203-221): filter in order
    if (lane == 0)
        for (uint32_t t = 0; t < sp_num; ++t) {
            const uint32_t v = sp_of(t);
            s.hash[(uint32_t)(hash64_2(v) & (H - 1))] = v;
        }
    for (uint32_t t = lane; t < sp_num; t += WAVE) {
        const uint32_t v = sp_of(t);
        s.cand[t] = make_key(dist_of(v), v);
    }
    wave_sync();
    wave_rank_sort(s.cand, sp_num, Fp, lane);
    wave_sync();
    for (uint32_t t = lane; t < sp_num; t += WAVE)
        s.unv[t] = Fp[t];
    uint32_t fs = sp_num, remain = sp_num, offset = 0, cand_n = 0, vs = 0;
    uint64_t nvis = 0, cmps = sp_num, nins = 0;
    wave_sync();

    while (remain > offset && nvis < limit) {
        const uint64_t cur = s.unv[offset];
        const uint32_t cid = key_id(cur);
        // 1) neighbour filter, then 2) distances (beamSearch.h:71-93)
        for (uint32_t j0 = 0; j0 < lim; j0 += WAVE) {
            const uint32_t j = j0 + lane;
            const uint32_t av = j < lim ? g.nbr[(uint64_t)cid * R + j] : NBPG_SENT;
            const bool valid = av != NBPG_SENT;
            const bool seen = filter_chunk(s.hash, H, av, valid, lane);
            const bool fresh = valid && !seen && av != self;
            const uint64_t um = ballot(fresh);
            if (fresh)
                s.cand[cand_n + prefix_count(um)] = make_key(dist_of(av), av);
            const uint32_t nf = (uint32_t)__popcll(um);
            cand_n += nf;
            cmps += nf;
            wave_sync();
        }
        // 3) sorted insertion of `cur` into visited (beamSearch.h:95-117)
        uint32_t c = 0;
        for (uint32_t i = lane; i < vs; i += WAVE)
            c += s.vis[i] <= cur;
        const uint32_t pos = wave_sum(c);
        ++nvis;
        const uint32_t nvs = vs < maxv ? vs + 1 : vs;
        if (pos + 1 < nvs) {
            for (int32_t top = (int32_t)nvs - 2; top >= (int32_t)pos; top -= WAVE) {
                const int32_t idx = top - (int32_t)lane;
                uint64_t v = 0;
                if (idx >= (int32_t)pos)
                    v = s.vis[idx];
                wave_sync();
                if (idx >= (int32_t)pos)
                    s.vis[idx + 1] = v;
                wave_sync();
            }
        }
        nins += pos < nvs;
        if (pos < nvs && lane == 0)
            s.vis[pos] = cur;
        vs = nvs;
        wave_sync();
        // 4) delayed merge (beamSearch.h:119-125)
        if (cand_n == 0 || (limit >= 2ull * L && cand_n < L / 8 && offset + 1 < remain)) {
            ++offset;
            continue;
        }
        offset = 0;
        // sort + unique by id (beamSearch.h:128-130)
        wave_rank_sort(s.cand, cand_n, s.cs, lane);
        wave_sync();
        uint32_t cu = 0;
        for (uint32_t i0 = 0; i0 < cand_n; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const uint64_t v = i < cand_n ? s.cs[i] : 0;
            const bool keep = i < cand_n && (i == 0 || key_id(s.cs[i - 1]) != key_id(v));
            const uint64_t km = ballot(keep);
            if (keep)
                s.cand[cu + prefix_count(km)] = v;
            cu += (uint32_t)__popcll(km);
        }
        wave_sync();
        // set_union(frontier, candidates) capped at L (beamSearch.h:132-138)
        uint32_t nb = 0;
        for (uint32_t i0 = 0; i0 < cu; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const uint64_t v = i < cu ? s.cand[i] : 0;
            const bool keep = i < cu && !contains(Fp, fs, v);
            const uint64_t km = ballot(keep);
            if (keep)
                s.cs[nb + prefix_count(km)] = v;
            nb += (uint32_t)__popcll(km);
        }
        wave_sync();
        for (uint32_t i = lane; i < fs; i += WAVE) {
            const uint64_t v = Fp[i];
            const uint32_t p = i + count_less(s.cs, nb, v);
            if (p < L)
                NFp[p] = v;
        }
        for (uint32_t j = lane; j < nb; j += WAVE) {
            const uint64_t v = s.cs[j];
            const uint32_t p = j + count_less(Fp, fs, v);
            if (p < L)
                NFp[p] = v;
        }
        uint32_t nfs = min(L, fs + nb);
        wave_sync();
        // cut at cut * d(k-th) (beamSearch.h:140-149)
        if (k > 0 && nfs > k) {
            const float thr = (float)(cut * (double)key_dist(NFp[k]));
            const uint64_t tk = make_key(thr, 0u);
            uint32_t cc = 0;
            for (uint32_t i = lane; i < nfs; i += WAVE)
                cc += NFp[i] <= tk;
            const uint32_t ub = wave_sum(cc);
            nfs = max(ub, fs);
        }
        {
            uint64_t *t = Fp;
            Fp = NFp;
            NFp = t;
        }
        fs = nfs;
        cand_n = 0;
        // unvisited = frontier[0:min(fs,L)] \ visited (beamSearch.h:155-160)
        const uint32_t fl = min(fs, L);
        uint32_t r = 0;
        for (uint32_t i0 = 0; i0 < fl; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const uint64_t v = i < fl ? Fp[i] : 0;
            const bool keep = i < fl && !contains(s.vis, vs, v);
            const uint64_t km = ballot(keep);
            if (keep)
                s.unv[r + prefix_count(km)] = v;
            r += (uint32_t)__popcll(km);
        }
        remain = r;
        wave_sync();
    }
    RefOut o;
    o.F = Fp;
    o.fs = fs;
    o.vs = vs;
    o.nvis = nvis;
    o.cmps = cmps;
    o.nins = nins;
    return o;
}

} // namespace nbpg
