// prune_mfma.h -- Prune::limitedSimilarityPrune (prune.h:158-209) with the
// candidate x candidate distances on int8 matrix cores, one wavefront per
// centre node.
//
// The candidates (up to 256, sorted and deduplicated by
// wave_sort_candidates) are held in vector registers as the operands of
// v_mfma_i32_16x16x64_i8: candidate 16 t + (lane % 16) of tile t, its
// values 64 s + 16 (lane / 16) .. + 16 in k-step s.  The same registers are
// the A operand (16 rows) and the B operand (16 columns) of a tile pair, so
// one product gives the 256 dot products <r, j> of a 16 x 16 block.
//
// Values enter the matrix core as signed bytes: u8 rows as x - 128 (x ^ 0x80),
// i8 rows as they are, packed 4-bit rows as one byte per nibble (0..15, the
// reference's unsigned reading of both 4-bit types, point.h:77-120).  Per
// candidate the self term c(x) is |x'|^2 (squared L2: translation leaves
// distances unchanged) or -A(x) with <x,y> = <x',y'> + A(x) + A(y) (inner
// product, A(x) = 128 sum(x') + 8192 bytes for u8, 0 otherwise), so the
// reference's integer distance value is
//     v(r, j) = c(r) + c(j) - g <r', j'>,   g = 2 (L2) or 1 (inner product),
// and its float distance is (float)(v * scale) (scale 256 for 4-bit rows),
// exactly what dist_from gives, since |v * scale| < 2^24 on this path.
//
// The alpha test alpha * (double)d(r, j) <= (double)d(p, j) (prune.h:227-230)
// is monotone in v, so per candidate j it is v <= T(j) for the largest such
// integer T(j), found once per candidate in double arithmetic; per element of
// a block the test is then one multiply-add and one compare, and the
// compare's lane mask is the block's row of prune bits.  With the angle test
// (cos_angle > 1e-5, prune.h:232-238) the block's elements evaluate both
// tests in the reference's own float and double steps instead.
//
// The greedy scan goes through the candidates 16 rows at a time (see
// wave_prune_mfma).  Selections and survivors are exactly those of the
// reference's scan (and of wave_prune).
#pragma once
#include <float.h>
#include <math.h>

#include "prune_core.h"

namespace nbpg {

typedef int pm_v4i __attribute__((ext_vector_type(4)));

#define PM_TILES 16
#define PM_MAXC (16 * PM_TILES) // candidates per node on this path

template <int DT, int NCH> struct prune_mfma_shape {
    static constexpr bool nib = DT == NBPG_U4X2 || DT == NBPG_I4X2;
    static constexpr int vals = nib ? 32 * NCH : 16 * NCH; // values per row
    static constexpr int KS = (vals + 63) / 64;            // k-steps of 64
    static constexpr bool ok = KS <= 2;
};

// per-wavefront LDS of the scan
struct PruneMfmaLds {
    int32_t ct[PM_MAXC]; // self term c(j)
    int32_t u[PM_MAXC];  // c(j) - T(j); with the angle test: d(p, j) as float bits
};

// fragment of candidate row `row` for k-step s and lane group kq < 4, as
// signed bytes (see above); zero past the row's end (rows that fill every
// k-step need no test)
template <int DT, int NCH>
__device__ __forceinline__ pm_v4i prune_frag(const uint8_t *__restrict__ row, int s, uint32_t kq)
{
    pm_v4i v = {0, 0, 0, 0};
    if (prune_mfma_shape<DT, NCH>::nib) {
        const uint32_t off = 32 * s + 8 * kq; // 8 packed bytes = 16 nibbles
        if (16 * NCH >= 32 * prune_mfma_shape<DT, NCH>::KS || off < 16 * NCH) {
            const uint2 u = *reinterpret_cast<const uint2 *>(row + off);
            v = pm_v4i{(int)(u.x & 0x0F0F0F0Fu), (int)((u.x >> 4) & 0x0F0F0F0Fu),
                       (int)(u.y & 0x0F0F0F0Fu), (int)((u.y >> 4) & 0x0F0F0F0Fu)};
        }
    } else {
        const uint32_t off = 64 * s + 16 * kq;
        if (16 * NCH >= 64 * prune_mfma_shape<DT, NCH>::KS || off < 16 * NCH) {
            const uint4 u = *reinterpret_cast<const uint4 *>(row + off);
            const uint32_t xm = DT == NBPG_U8 ? 0x80808080u : 0u;
            v = pm_v4i{(int)(u.x ^ xm), (int)(u.y ^ xm), (int)(u.z ^ xm), (int)(u.w ^ xm)};
        }
    }
    return v;
}

// largest integer v in [-2^24, 2^24] with alpha * (double)(v * scale) <= (double)dpj,
// -2^24 - 1 when there is none (also for a NaN distance); alpha > 1e-5.
// alpha * x is monotone in x under rounding, so the values that pass form a
// prefix; the quotient's rounding error is far below one unit, so a few
// steps from floor(dpj / (alpha * scale)) reach the exact end of the prefix.
__device__ __forceinline__ int32_t prune_bound(float dpj, double alpha, double scale)
{
    const int64_t LO = -(1ll << 24), HI = 1ll << 24;
    const double K = (double)dpj;
    if (!(K == K))
        return (int32_t)(LO - 1);
    const double g = floor(K / (alpha * scale));
    int64_t v = g < (double)LO ? LO - 1 : (g > (double)HI ? HI : (int64_t)g);
    for (int it = 0; it < 4 && v < HI && alpha * ((double)(v + 1) * scale) <= K; ++it)
        ++v;
    for (int it = 0; it < 4 && v >= LO && !(alpha * ((double)v * scale) <= K); ++it)
        --v;
    return (int32_t)v;
}

// the 16 columns of a 64-lane compare mask (lane 16 kq + c: column c) that
// are set in any of the four lane groups
__device__ __forceinline__ uint32_t pm_fold(uint64_t b)
{
    const uint32_t x = (uint32_t)b | (uint32_t)(b >> 32);
    return (x | (x >> 16)) & 0xFFFFu;
}

// limitedSimilarityPrune over sorted keys[0:m), m <= PM_MAXC, for centre p
// with alpha > 1e-5 or cos_angle > 1e-5; writes up to R survivors to
// out_ids/out_d and returns their count (distance evaluations are not
// counted on this path).
//
// Per tile of 16 rows, in order: the diagonal block (the tile against
// itself) gives each row's prune bits over the tile's own later rows, since
// only earlier rows of the same tile can prune a row of it (earlier tiles
// have been applied); a scalar pass over the tile's live rows picks the
// selected ones from those bits.  The later column tiles then need only
// "pruned by some selected row of the tile": with y(r, j) = g <r', j'> - c(r)
// and w(j) = c(j) - T(j), the test is y >= w, and an unselected row's c(r) is
// replaced by 2^30 so that max over a lane's four rows >= w is the whole
// test for the lane's column; one compare and one ballot per block, folded
// over the four lane groups, clear the pruned columns' live bits.
template <int DT, int NCH, bool ANGLE>
__device__ uint32_t wave_prune_mfma(const uint64_t *keys, uint32_t m, uint32_t p, uint32_t R,
                                    double alpha, double cosa, const uint8_t *__restrict__ vecs,
                                    uint32_t stride, PruneMfmaLds &S, uint32_t *out_ids,
                                    float *out_d, uint32_t lane, uint32_t mips)
{
    using SH = prune_mfma_shape<DT, NCH>;
    constexpr int KS = SH::KS;
    constexpr int32_t FAR = 1 << 30;
    const double eps = 1e-5;
    const uint32_t col = lane % 16, kq = lane / 16;
    const uint32_t ntiles = (m + 15) / 16;
    const uint32_t sh = mips ? 0u : 1u; // g = 1 << sh
    const double scale = (double)dot_traits<DT>::scale;

    // candidate rows -> fragments (all loads issued before the first use).
    // Lanes past the last candidate load the centre's row instead: their
    // columns are never live and their rows never selected, so the values
    // are never read, and every lane's load is then unconditional.
    pm_v4i F[PM_TILES][KS];
#pragma unroll
    for (int t = 0; t < PM_TILES; ++t) {
        const uint32_t c = 16 * t + col;
        const uint8_t *row = vecs + (uint64_t)(c < m ? key_id(keys[c]) : p) * stride;
#pragma unroll
        for (int s = 0; s < KS; ++s)
            F[t][s] = prune_frag<DT, NCH>(row, s, kq);
    }
    // self terms: each candidate's four lane groups summed over lanes l ^ 16,
    // l ^ 32 (one copy per metric, so the products take no per-word select)
    auto self_terms = [&](auto IP) {
        constexpr bool ip = decltype(IP)::value;
#pragma unroll
        for (int t = 0; t < PM_TILES; ++t) {
            int acc = 0;
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int w = 0; w < 4; ++w)
                    acc = __builtin_amdgcn_sdot4(F[t][s][w], ip ? 0x01010101 : F[t][s][w], acc,
                                                 false);
            acc += __shfl_xor(acc, 16);
            acc += __shfl_xor(acc, 32);
            const int32_t ct =
                !ip ? acc : (DT == NBPG_U8 ? -(128 * acc + 8192 * (int32_t)(16 * NCH)) : 0);
            if (kq == 0)
                S.ct[16 * t + col] = ct;
        }
    };
    if (mips)
        self_terms(std::true_type());
    else
        self_terms(std::false_type());
    wave_sync();
    // per candidate: w(j) = c(j) - T(j) (or, with the angle test, d(p, j))
    for (uint32_t j = lane; j < m; j += WAVE) {
        const float dpj = key_dist(keys[j]);
        if (ANGLE)
            S.u[j] = __float_as_int(dpj);
        else
            S.u[j] = S.ct[j] - prune_bound(dpj, alpha, scale);
    }
    // live candidates, one 16-bit field per tile (wave-uniform)
    uint32_t live[PM_TILES];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t j = 64 * w + lane;
        const uint32_t id = j < m ? key_id(keys[j]) : NBPG_SENT;
        const uint64_t b = ballot(j < m && id != p && id != NBPG_SENT);
#pragma unroll
        for (int f = 0; f < 4; ++f)
            live[4 * w + f] = (uint32_t)(b >> (16 * f)) & 0xFFFFu;
    }
    wave_sync();
    // this lane's column term of every tile, read once
    int32_t wcol[PM_TILES];
#pragma unroll
    for (int tc = 0; tc < PM_TILES; ++tc)
        wcol[tc] = S.u[16 * tc + col];
    // prune bits of row tile tr (rows 4 kq + i) against column tile tc; `ncr`
    // holds -c(r) per row (FAR rows never prune)
    auto block = [&](int tr, int tc, const pm_v4i &ncr, const pm_v4i &dpr, bool (&pr)[4],
                     int32_t &ymax, int32_t &wmax) {
        pm_v4i acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KS; ++s)
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(F[tr][s], F[tc][s], acc, 0, 0, 0);
        // this lane's column: w(j) (or d(p, j) bits and c(j))
        const int32_t wj = ANGLE ? S.u[16 * tc + col] : wcol[tc];
        int32_t y[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            y[i] = (int32_t)(((uint32_t)acc[i] << sh) + (uint32_t)ncr[i]);
        if (ANGLE) {
            const int32_t cj = S.ct[16 * tc + col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // v = c(r) + c(j) - g <r', j'>; a FAR row never prunes
                const int32_t v = cj - y[i];
                const float dsp = (float)(v * (int32_t)dot_traits<DT>::scale);
                const float dpj = __int_as_float(wj), dr = __int_as_float(dpr[i]);
                pr[i] = ncr[i] != -FAR &&
                        ((alpha > eps && alpha * (double)dsp <= (double)dpj) ||
                         (cosa > eps && (double)cosine_at_p(dr, dpj, dsp) > cosa));
            }
        } else {
            ymax = max(max(y[0], y[1]), max(y[2], y[3]));
#pragma unroll
            for (int i = 0; i < 4; ++i)
                pr[i] = y[i] >= wj;
        }
        wmax = wj;
    };

    uint32_t sel = 0;
#pragma unroll
    for (int tr = 0; tr < PM_TILES; ++tr) {
        if ((uint32_t)tr < ntiles && sel < R && live[tr] != 0) {
            const pm_v4i crow = *reinterpret_cast<const pm_v4i *>(&S.ct[16 * tr + 4 * kq]);
            pm_v4i dpr = {0, 0, 0, 0};
            if (ANGLE)
                dpr = *reinterpret_cast<const pm_v4i *>(&S.u[16 * tr + 4 * kq]);
            const pm_v4i ncr = -crow;
            // the diagonal block: row 4 kq + i's bits over the tile in b[i]
            bool pr[4];
            int32_t ymax, wd;
            block(tr, tr, ncr, dpr, pr, ymax, wd);
            uint64_t b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                b[i] = ballot(pr[i]);
            // row r's bits over the tile in lane r (r < 16), read back by lane
            // index in the scan below
            uint32_t rmask;
            {
                const uint32_t r = lane & 15u, ri = r & 3u;
                const uint64_t b01 = (ri & 1u) ? b[1] : b[0], b23 = (ri & 1u) ? b[3] : b[2];
                const uint64_t bb = (ri & 2u) ? b23 : b01;
                rmask = (uint32_t)(bb >> (16 * (r >> 2)));
            }
            // the tile's live rows in order (prune.h:174-206)
            uint32_t todo = live[tr], ssel = 0;
            const uint32_t sel0 = sel, room = R - sel; // room > 0
            while (todo != 0) {
                const uint32_t rl = (uint32_t)__builtin_ctz(todo);
                ssel |= 1u << rl;
                todo &= ~(uint32_t)__builtin_amdgcn_readlane((int)rmask, (int)rl);
                todo &= ~((2u << rl) - 1u); // rows up to rl are done
                if (++sel - sel0 == room)
                    break;
            }
            live[tr] = 0;
            // the selected rows, in order, to the output
            if (lane < 16 && ((ssel >> lane) & 1u)) {
                const uint32_t pos = sel0 + (uint32_t)__popc(ssel & ((1u << lane) - 1u));
                const uint64_t kr = keys[16 * tr + lane];
                out_ids[pos] = key_id(kr);
                if (out_d)
                    out_d[pos] = key_dist(kr);
            }
            if (ssel != 0 && sel < R) {
                // rows not selected stop pruning
                pm_v4i ncm;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    ncm[i] = ((ssel >> (4 * kq + i)) & 1u) ? ncr[i] : -FAR;
                auto column = [&](int tc) {
                    if (live[tc] == 0) // every candidate of the tile is pruned already
                        return;
                    bool pc[4];
                    int32_t ym = 0, wj = 0;
                    block(tr, tc, ncm, dpr, pc, ym, wj);
                    const bool any = ANGLE ? (pc[0] || pc[1] || pc[2] || pc[3]) : ym >= wj;
                    live[tc] &= ~pm_fold(ballot(any));
                };
                if (ntiles == PM_TILES) {
#pragma unroll
                    for (int tc = tr + 1; tc < PM_TILES; ++tc)
                        column(tc);
                } else {
#pragma unroll
                    for (int tc = tr + 1; tc < PM_TILES; ++tc)
                        if ((uint32_t)tc < ntiles)
                            column(tc);
                }
            }
        }
    }
    return sel;
}

} // namespace nbpg
