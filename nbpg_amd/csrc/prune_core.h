// prune_core.h -- one wavefront's Prune::sort_candidates (prune.h:49-74)
// and Prune::limitedSimilarityPrune (prune.h:158-209), shared by the
// batched prune kernel and the device build's insert kernel.
#pragma once
#include <float.h>

#include "internal.h"

namespace nbpg {

// ascending bitonic sort of a[0:n2) (n2 a power of two) by one wavefront
__device__ __forceinline__ void wave_bitonic(uint64_t *a, uint32_t n2, uint32_t lane)
{
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n2; i += WAVE) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = a[i], y = a[p];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[p] = x;
                    }
                }
            }
            wave_sync();
        }
    }
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n)
{
    uint32_t n2 = 1;
    while (n2 < n)
        n2 <<= 1;
    return n2;
}

// keys[0:n) hold (dist, id) keys with the centre p already keyed at FLT_MAX
// (biased_less with dir_bias 0); keys must have room for pow2_at_least(n).
// Sorts, collapses equal ids (dedupe_by_index: equal ids carry equal keys),
// drops trailing centre entries.  Returns the new length.
__device__ __forceinline__ uint32_t wave_sort_candidates(uint64_t *keys, uint32_t n, uint32_t p,
                                                         uint32_t lane)
{
    // candidates that arrive in (dist, id) order (a beam, a sorted list)
    // skip the sort: one pass over adjacent pairs decides
    bool unsorted = false;
    for (uint32_t i = lane; i + 1 < n; i += WAVE)
        unsorted |= keys[i] > keys[i + 1];
    if (ballot(unsorted)) {
        const uint32_t n2 = pow2_at_least(n);
        for (uint32_t i = n + lane; i < n2; i += WAVE)
            keys[i] = ~0ull;
        wave_sync();
        wave_bitonic(keys, n2, lane);
    }
    uint32_t m = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += WAVE) {
        const uint32_t i = i0 + lane;
        const uint64_t v = i < n ? keys[i] : 0;
        const bool keep = i < n && (i == 0 || key_id(keys[i - 1]) != key_id(v));
        const uint64_t km = ballot(keep);
        wave_sync();
        if (keep)
            keys[m + prefix_count(km)] = v;
        m += (uint32_t)__popcll(km);
        wave_sync();
    }
    while (m > 0 && key_id(keys[m - 1]) == p)
        --m;
    return m;
}

__device__ __forceinline__ float cosine_at_p(float d_ps, float d_pp, float d_sp)
{
    // prune.h:218-225
    const float b = sqrtf(d_pp), c = sqrtf(d_ps);
    if (b == 0.f || c == 0.f)
        return 1.f;
    return (d_pp + d_ps - d_sp) / (2.f * b * c);
}

// limitedSimilarityPrune over sorted keys[0:m) for centre p; writes up to R
// survivors to out_ids/out_d (lane 0 stores); returns the survivor count.
// `comps` accumulates the distance evaluations (the reference's return).
template <int DT, int NCH>
__device__ uint32_t wave_prune(uint64_t *keys, uint32_t m, uint32_t p, uint32_t R, double alpha,
                               double cosa, const uint8_t *__restrict__ vecs, uint32_t stride,
                               const uint32_t *__restrict__ norms, uint32_t *out_ids, float *out_d,
                               uint32_t &comps, uint32_t lane, uint32_t mips = 0)
{
    const double eps = 1e-5;
    uint32_t sel = 0;
    if (alpha <= eps && cosa <= eps) { // prune.h:164-169
        sel = min(m, R);
        for (uint32_t i = lane; i < sel; i += WAVE) {
            out_ids[i] = key_id(keys[i]);
            if (out_d)
                out_d[i] = key_dist(keys[i]);
        }
        return sel;
    }
    uint32_t scan = 0;
    while (sel < R && scan < m) {
        const uint64_t ks = keys[scan];
        const uint32_t ps = key_id(ks);
        ++scan;
        if (ps == p || ps == NBPG_SENT)
            continue;
        const float d_ps = key_dist(ks);
        if (lane == 0) {
            out_ids[sel] = ps;
            if (out_d)
                out_d[sel] = d_ps;
        }
        ++sel;
        uint4 sv[NCH];
        const uint4 *sr = reinterpret_cast<const uint4 *>(vecs + (uint64_t)ps * stride);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            sv[c] = sr[c];
        const uint32_t ns = norms[ps];
        uint32_t ss[4 * NCH];
        to_uniform<NCH>(sv, ss);
        for (uint32_t i0 = scan; i0 < m; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const uint64_t kv = i < m ? keys[i] : ~0ull;
            const uint32_t pp = key_id(kv);
            const bool live = i < m && pp != NBPG_SENT;
            bool prune = false;
            if (live) {
                const uint4 *r = reinterpret_cast<const uint4 *>(vecs + (uint64_t)pp * stride);
                const float d_sp = dist_from<DT>(norms[pp], ns, dot_row<DT, NCH>(r, ss), mips);
                const float d_pp = key_dist(kv);
                // prune.h:227-238, alpha test in double
                if (alpha > eps && alpha * (double)d_sp <= (double)d_pp)
                    prune = true;
                else if (cosa > eps && (double)cosine_at_p(d_ps, d_pp, d_sp) > cosa)
                    prune = true;
            }
            comps += (uint32_t)__popcll(ballot(live));
            if (prune)
                keys[i] = (kv & 0xFFFFFFFF00000000ull) | NBPG_SENT;
        }
        wave_sync();
    }
    return sel;
}

// The same scan with the candidates' rows, keys and live flags held in
// registers: candidate i = 64 k + lane lives in slot k of lane i % 64,
// loaded once.  The next selected p* is the first live candidate past the
// previous one (one ballot per slot, not one step per pruned entry), its row
// comes from the lane that holds it, and the pass over the later candidates
// prunes in registers.  Needs m <= 64 * SLOTS.  Pairs, tests, distance
// counts and survivors are exactly those of wave_prune.
template <int DT, int NCH, int SLOTS>
__device__ uint32_t wave_prune_regs(uint64_t *keys, uint32_t m, uint32_t p, uint32_t R, double alpha,
                                    double cosa, const uint8_t *__restrict__ vecs, uint32_t stride,
                                    const uint32_t *__restrict__ norms, uint32_t *out_ids, float *out_d,
                                    uint32_t &comps, uint32_t lane, uint32_t mips = 0)
{
    const double eps = 1e-5;
    uint4 rows[SLOTS][NCH];
    uint32_t nrm[SLOTS];
    float kd[SLOTS]; // d(p, candidate)
    bool live[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const uint32_t i = WAVE * k + lane;
        const uint64_t kv = i < m ? keys[i] : ~0ull;
        const uint32_t id = key_id(kv);
        live[k] = i < m && id != NBPG_SENT && id != p;
        kd[k] = key_dist(kv);
        nrm[k] = 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            rows[k][c] = make_uint4(0, 0, 0, 0);
        if (i < m && id != NBPG_SENT) {
            const uint4 *r = reinterpret_cast<const uint4 *>(vecs + (uint64_t)id * stride);
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                rows[k][c] = r[c];
            nrm[k] = norms[id];
        }
    }
    uint32_t sel = 0, scan = 0;
    while (sel < R) {
        // the first live candidate at or past `scan`
        uint32_t si = 0xFFFFFFFFu;
#pragma unroll
        for (int k = SLOTS - 1; k >= 0; --k) {
            const uint64_t b = ballot(live[k] && WAVE * (uint32_t)k + lane >= scan);
            if (b)
                si = WAVE * (uint32_t)k + (uint32_t)__builtin_ctzll(b);
        }
        if (si == 0xFFFFFFFFu)
            break;
        const uint32_t sk = si / WAVE, sl = si % WAVE;
        uint32_t ss[4 * NCH];
        uint32_t ns = 0, ps = 0;
        float d_ps = 0.f;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k)
            if ((uint32_t)k == sk) {
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    ss[4 * c + 0] = __builtin_amdgcn_readlane(rows[k][c].x, sl);
                    ss[4 * c + 1] = __builtin_amdgcn_readlane(rows[k][c].y, sl);
                    ss[4 * c + 2] = __builtin_amdgcn_readlane(rows[k][c].z, sl);
                    ss[4 * c + 3] = __builtin_amdgcn_readlane(rows[k][c].w, sl);
                }
                ns = __builtin_amdgcn_readlane(nrm[k], sl);
                d_ps = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(kd[k]), sl));
            }
        ps = key_id(keys[si]);
        if (lane == 0) {
            out_ids[sel] = ps;
            if (out_d)
                out_d[sel] = d_ps;
        }
        ++sel;
        scan = si + 1;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if (WAVE * (uint32_t)k + WAVE <= scan || WAVE * (uint32_t)k >= m)
                continue; // uniform: the slot holds no candidate past p*
            const uint32_t i = WAVE * k + lane;
            const bool cand = live[k] && i >= scan;
            if (cand) {
                uint32_t a0 = 0, a1 = 0;
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    a0 = dot_traits<DT>::dot(rows[k][c].x, ss[4 * c + 0], a0);
                    a1 = dot_traits<DT>::dot(rows[k][c].y, ss[4 * c + 1], a1);
                    a0 = dot_traits<DT>::dot(rows[k][c].z, ss[4 * c + 2], a0);
                    a1 = dot_traits<DT>::dot(rows[k][c].w, ss[4 * c + 3], a1);
                }
                const float d_sp = dist_from<DT>(nrm[k], ns, a0 + a1, mips);
                const float d_pp = kd[k];
                // prune.h:227-238, alpha test in double
                if (alpha > eps && alpha * (double)d_sp <= (double)d_pp)
                    live[k] = false;
                else if (cosa > eps && (double)cosine_at_p(d_ps, d_pp, d_sp) > cosa)
                    live[k] = false;
            }
            comps += (uint32_t)__popcll(ballot(cand));
        }
    }
    return sel;
}

} // namespace nbpg
