/*
 * nbpg_oracle.h -- CPU restatement of the reference NBPG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; nothing in nbpg_amd/ links,
 * imports or calls it.
 *
 * Every function restates one reference routine, cited as file:line
 * relative to the reference tree (godpgf/NBPG).  The restatement is plain
 * C11 + OpenMP, written from the reference's behaviour, not its text.
 */
#ifndef NBPG_ORACLE_H
#define NBPG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types, selected by file extension in the reference
 * (Vamana.cpp:292-311, CompGT.cpp:152-163) */
enum { NBPGO_U8 = 0, NBPGO_I8 = 1, NBPGO_U4X2 = 2, NBPGO_I4X2 = 3, NBPGO_F32 = 4 };
/* OR into a dtype: the inner-product metric (DistanceMetric::MIPS, point.h:20-24) */
enum { NBPGO_MIPS = 0x100 };

/* (id, dist) record; order = dist asc, id asc (query_params.h:97-119) */
typedef struct {
    uint32_t id;
    float dist;
} nbpgo_iddist;

/* a vector set: rows of `dims` elements (bytes for the packed 4-bit types),
 * `stride` bytes apart (point_range.h:48-57 pads to 64 B; any stride works) */
typedef struct {
    const uint8_t *data;
    uint64_t stride;
    uint64_t n;
    uint32_t dims;
    int32_t dtype;
} nbpgo_points;

/* adjacency in the reference's in-memory layout: row i is
 * [deg, id_0 .. id_{R-1}], R+1 u32 per row (graph.h:142-161, 381-391) */
typedef struct {
    uint32_t *rows;
    uint64_t n;
    uint32_t R;
} nbpgo_graph;

/* QueryParams (query_params.h:12-42), pure-quantized fields only */
typedef struct {
    uint32_t k;
    double cut;
    uint32_t L;            /* beamSize */
    uint64_t limit;        /* max expansions */
    uint32_t degree_limit; /* neighbours expanded per node */
    uint64_t max_visited;  /* visited-list cap: 8L at query time (beamSearch.h:279,290) */
} nbpgo_qparams;

/* distances (point.h:27-156) */
float nbpgo_distance(int dtype, const void *a, const void *b, uint32_t dims);
uint64_t nbpgo_hash64_2(uint64_t x);
uint32_t nbpgo_hash_filter_size(uint32_t beam);

/* CompGT (CompGT.cpp:29-134).  tie_mode 0 = the reference's eviction rule,
 * 1 = exact lexicographic (dist, id) top-k.  Needs n >= k. */
int nbpgo_compgt(int dtype, const void *base, uint64_t n, const void *queries, uint64_t nq,
                 uint32_t dims, uint32_t k, int tie_mode, uint32_t *out_ids, float *out_dists,
                 int nthreads);

/* Prune::sort_candidates with centre p, dir_bias 0 (prune.h:49-74) */
uint64_t nbpgo_sort_candidates(uint32_t p, nbpgo_iddist *cand, uint64_t n);
/* Prune::limitedSimilarityPrune (prune.h:158-209); returns distance count */
uint64_t nbpgo_robust_prune(uint32_t p, nbpgo_iddist *cand, uint64_t *n, const nbpgo_points *P,
                            uint32_t R, double alpha, double cos_angle, uint64_t start_prune_id);

/* qsearchAll + filtered_beam_search, pure-quantized mode
 * (beamSearch.h:43-171, 183-303).  out_ids/out_dists are nq*k (slots past the
 * frontier size get id 0xFFFFFFFF, dist +inf); out_stats is nq*2 = {visited, dist_cmps}. */
int nbpgo_search(const nbpgo_graph *G, const nbpgo_points *base, const nbpgo_points *queries,
                 const uint32_t *start_points, uint32_t sp_num, const nbpgo_qparams *qp,
                 uint32_t *out_ids, float *out_dists, uint32_t *out_stats, int nthreads);

/* qsearchAll with raw rows beside the quantized ones (beamSearch.h:183-247):
 * rerank_factor > 0 rerank mode, 0 filtering mode.  Raw rows are distances
 * only (raw_base->n == base->n). */
int nbpgo_search_exact(const nbpgo_graph *G, const nbpgo_points *base,
                       const nbpgo_points *queries, const nbpgo_points *raw_base,
                       const nbpgo_points *raw_queries, const uint32_t *start_points,
                       uint32_t sp_num, const nbpgo_qparams *qp, uint32_t rerank_factor,
                       uint32_t *out_ids, float *out_dists, uint32_t *out_stats, int nthreads);

/* one insert-time search + prune for node `index` (nsw.h:122-175).
 * out: pruned (id, dist) list (<= R) and the beam's visited list. */
typedef struct {
    uint32_t L, VL, R;
    uint64_t limit;
} nbpgo_insert_params;

/* build_vamana (nsw.h:547-663) */
typedef struct {
    uint32_t R, L, VL, top_degree, num_split, num_passes;
    uint64_t limit;
    double min_alpha, max_alpha, min_cos_angle, max_cos_angle;
    int32_t dynamic_prune, rand_sp, repair_isolate;
    uint64_t sp_seed; /* seed of the start-point stream that replaces randomInt (utilities.h:55-63) */
} nbpgo_build_params;

int nbpgo_build(const nbpgo_points *P, nbpgo_graph *G, const nbpgo_build_params *bp,
                uint64_t *out_isolated, int nthreads);

/* libstdc++ std::shuffle(first, last, std::mt19937(seed)) over u32 */
void nbpgo_shuffle_mt19937(uint32_t *a, uint64_t n, uint32_t seed);
/* the seeded start-point stream used in place of randomInt */
uint64_t nbpgo_sp_stream(uint64_t seed, uint64_t counter);

/* DFS::repair_isolate_node(G, sp) (dfs.h:192-215, 248-253) */
uint64_t nbpgo_repair_isolated(nbpgo_graph *G, uint32_t sp);

/* matmul (matrix.h:86-97): y = bias + W x, sequential fp32 sums */
void nbpgo_matmul(const float *x, uint64_t n, uint32_t d_in, const float *W, const float *bias,
                  uint32_t d_out, float *y, int nthreads);
/* compress (matrix.h:99-127): y (n x d) -> codes (n x d/chunk bytes) */
void nbpgo_compress(const float *y, uint64_t n, uint32_t d, int is_signed, uint32_t chunk,
                    uint8_t *out, int nthreads);

/* calRecall with tie extension (check_nn_recall.h:126-167) */
double nbpgo_recall(const uint32_t *result_ids, uint32_t res_stride, uint64_t nq,
                    const uint32_t *gt_ids, uint32_t gt_k, const nbpgo_points *base,
                    const nbpgo_points *queries, uint32_t k);

#ifdef __cplusplus
}
#endif
#endif
