// search_args.h -- launch arguments and helpers shared by the search kernels.
#pragma once
#include "raw_rows.h"

namespace nbpg {

struct SearchArgs {
    const uint8_t *vecs;
    const uint32_t *norms;
    const uint32_t *nbr;
    uint64_t n;
    uint32_t stride, R;
    const uint8_t *queries;
    uint64_t qstride, nq;
    uint32_t qbytes;
    uint32_t mips; // inner-product metric
    // reference mode
    uint32_t sp[64]; // start points: reference mode (<= 32) or FAST seeds (<= 64)
    uint32_t sp_num;
    uint32_t k, L, dlim, maxv, H;
    double cut;
    uint64_t limit;
    uint8_t *gscratch;
    uint64_t state_bytes;
    // reference mode over full-precision rows (beamSearch.h:183-247):
    // exact = 1 rerank of frontier[0:min(k * rerank_factor, fs)], 2 filtering
    uint32_t exact, raw_dtype, raw_dims, rerank_factor;
    const uint8_t *raw_base, *raw_q;
    uint64_t raw_row, raw_qstride;
    // fast mode
    const uint32_t *Ls, *ks;
    uint32_t out_k, Lmax, W, flog2, entry, wave_bytes;
    uint32_t nseeds; // FAST: > 1 seeds in sp[] replace the single entry
    uint32_t qstart; // FAST: 1 = query q starts from out_ids[q * out_k] (nearest seed)
    uint32_t *counter;
    // outputs
    uint32_t *out_ids;
    float *out_dists;
    uint32_t *out_stats;
    uint32_t stats_w; // u32 per stats row: 2 {visited, dist_cmps} or 3 (+ adjacency entries read)
    // FAST: the expanded nodes in expansion order, vis_cap per query, ids
    // past the count NBPG_SENT (NULL: not kept)
    uint32_t *vis_out;
    uint32_t vis_cap;
    uint32_t vis_cmps; // 1: the list is re-scored afterwards, count it in dist_cmps
};

#define FAST_FLAG 0x80000000ull // bit 31 of the id word: entry expanded

// next query of a persistent launch (>= nq: none left).  A wavefront's first
// query is its own index in the grid, taken without an atomic: every fetch
// of one counter is serialised in the memory system (about 11 ns apart),
// which at launch would start the grid's waves one by one over tens of
// microseconds.  Later queries come from counter[0], offset by the grid's
// wave count.  counter[48] (another 128-byte line, so the exits do not
// contend with the hand-out) counts the wavefronts that found none, and the
// last of them returns both to zero, so consecutive launches on one counter
// (eager or replayed from a graph) need no clearing step between them.
// Counters start at zero when allocated.  `first` starts true per wave.
__device__ __forceinline__ uint32_t next_query(uint32_t *counter, uint64_t nq, uint32_t lane, bool &first)
{
    const uint32_t waves = gridDim.x * (blockDim.x / WAVE);
    uint32_t qi = 0;
    const bool from_grid = first;
    first = false;
    if (from_grid)
        qi = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE);
    if (lane == 0) {
        if (!from_grid)
            qi = waves + atomicAdd(counter, 1u);
        // the exit count is taken after the hand-out atomic has returned (its
        // value decides the branch), so no fence is needed -- and a
        // device-scope fence here would write back and invalidate this XCD's L2
        if (qi >= nq) {
            if (atomicAdd(counter + 48, 1u) == waves - 1) {
                atomicExch(counter, 0u);
                atomicExch(counter + 48, 0u);
            }
        }
    }
    return __builtin_amdgcn_readfirstlane(qi);
}

__device__ __forceinline__ uint64_t unflag(uint64_t k) { return k & ~FAST_FLAG; }

__device__ __forceinline__ uint32_t fast_slot(uint32_t id, uint32_t flog2)
{
    return (id * 0x9E3779B1u) >> (32 - flog2);
}

// FAST-mode launch (search.hip); counter is the launch's own work counter
// seeds: nseeds > 1 distinct start nodes (host array) instead of entry
// stats_w: 2 or 3 u32 per stats row (3 adds the adjacency entries read)
// vis_out: also the first vis_cap expanded nodes of each query (d_ids may
// then be NULL); L <= 256 only
int search_fast(const nbpg_index *ix, const void *d_queries, uint64_t nq, uint64_t qstride,
                uint32_t entry, const uint32_t *seeds, uint32_t nseeds, uint32_t L, uint32_t k,
                const uint32_t *d_L, const uint32_t *d_k, uint32_t Lmax, uint32_t out_k,
                uint64_t limit, const nbpg_fast_config *cfg, uint32_t *d_ids, float *d_dists,
                uint32_t *d_stats, uint32_t *counter, hipStream_t stream, uint32_t stats_w = 2,
                uint32_t *vis_out = nullptr, uint32_t vis_cap = 0, bool vis_cmps = false,
                bool qstart = false);

int launch_search_reg(int dtype, uint32_t nch, uint32_t ipl, uint32_t E, bool split, dim3 grid,
                      dim3 block, uint32_t lds, hipStream_t stream, const SearchArgs &a);

} // namespace nbpg
