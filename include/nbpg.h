/*
 * nbpg.h -- C-ABI of the MI355X (gfx950) NBPG hot path.
 *
 * One shared library (nbpg_amd/libnbpg_amd.so), HIP only.  Every compute
 * entry point takes DEVICE pointers for its inputs and outputs and an
 * optional hipStream_t (passed as void*, NULL = the legacy default stream),
 * and returns an int status (NBPG_OK or a negative code; never aborts).
 * nbpg_last_error() gives the message of the calling thread's last failure.
 * Searches on one index (nbpg_search, nbpg_search_uneven, graph replays)
 * may run at the same time on different streams: each launch takes the
 * next of the index's 64 work counters (so at most 64 may be in flight on
 * the device at once), and a search graph owns one.  Calls that use the
 * index's workspace -- nbpg_search_rerank, NBPG_SEARCH_REF when a query's
 * state is over 16 KB (L above about 64 at R 64) and the call has more
 * queries than LDS holds at once, prune, build, repair -- must not overlap
 * other calls on the index (issue them on one stream, or synchronise
 * between streams); a server (nbpg_server_*) keeps counters of its own.
 *
 * Citations are file:line in the reference tree (godpgf/NBPG), naming the
 * interface each entry point stands in for.
 */
#ifndef NBPG_H
#define NBPG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference aborts instead: beamSearch.h:268-273,
 *      graph.h:381-387, query_params.h:296-300) ---- */
#define NBPG_OK 0
#define NBPG_EINVAL -1   /* bad argument: k > L, id out of range, ... */
#define NBPG_ENOMEM -2   /* device allocation failed */
#define NBPG_EHIP -3     /* a HIP runtime call failed */
#define NBPG_EUNSUP -4   /* shape the kernels do not support */

/* ---- element types; the reference picks them by file extension
 *      (Vamana.cpp:292-311, CompGT.cpp:152-163).  dims counts elements,
 *      i.e. BYTES for the packed 4-bit types (point.h:167-169). ---- */
#define NBPG_U8 0
#define NBPG_I8 1
#define NBPG_U4X2 2 /* .u4bin: two 4-bit codes per byte */
#define NBPG_I4X2 3 /* .i4bin: distance identical to U4X2 (point.h:107-120) */
#define NBPG_F32 4  /* .fbin / .bin: PCA input and raw rerank rows (matrix.h:27-57) */

/* ---- distance: DistanceMetric (point.h:20-24), chosen by --use_mips
 *      (Vamana.cpp:76, CompGT.cpp:147) ---- */
#define NBPG_METRIC_L2 0   /* squared Euclidean (the default) */
#define NBPG_METRIC_MIPS 1 /* negated inner product, mips_distance (point.h:46-156) */

/* ---- search modes ---- */
#define NBPG_SEARCH_REF 0  /* one query per workgroup, reference-exact IDs */
#define NBPG_SEARCH_FAST 1 /* persistent, one wavefront per query, LDS filter + beam */
/* OR into nbpg_search's mode: write the whole frontier (up to L per query,
 * output stride L) instead of its first k; the reference mode still cuts
 * with k.  This is the candidate list the rerank step reads. */
#define NBPG_SEARCH_FRONTIER 0x10
/* OR into nbpg_search's NBPG_SEARCH_FAST mode: stats rows get a third u32,
 * the adjacency entries the query's expanded nodes hold (the sum of their
 * degrees), for byte accounting; d_out_stats then holds nq*3 u32. */
#define NBPG_SEARCH_EDGES 0x20
/* OR into nbpg_search's NBPG_SEARCH_FAST mode: write each query's visited
 * list -- its expanded nodes in expansion order, the reference's
 * bsCell.visited (beamSearch.h:43-171) as a set -- instead of its top k:
 * NBPG_VISITED_CAP(L) ids per query (output stride), 0xFFFFFFFF past the
 * count; d_out_dists is not written.  L <= 256. */
#define NBPG_SEARCH_VISITED 0x40
/* OR into nbpg_search's NBPG_SEARCH_FAST mode: every query starts from the
 * nearest of the index's seed rows (nbpg_index_set_seeds) -- the per-query
 * start points the reference's driver asks get_sp(i) for
 * (beamSearch.h:257-303, check_nn_recall.h:282-297) -- instead of the host
 * start points (still checked, not used).  A pass before the search finds
 * each query's nearest seed from exact integer distances on int8 matrix
 * cores (ties to the lower id) and leaves its id in the query's first
 * d_out_ids entry, where the search reads it before writing its results
 * (so d_out_ids is required).  u8 / i8 rows of a multiple of 16 bytes,
 * 16-byte aligned query rows, L <= 256; not with NBPG_SEARCH_VISITED. */
#define NBPG_SEARCH_SEED_NEAREST 0x80
#define NBPG_MAX_SEEDS 4096
/* visited-list capacity per query: the reference's 8L (beamSearch.h:279),
 * at most 1024 */
#define NBPG_VISITED_CAP(L) ((8u * (L)) < 1024u ? (8u * (L)) : 1024u)

/* ---- CompGT modes ---- */
#define NBPG_GT_PLAIN 0 /* VALU dot4, one query tile per workgroup */
#define NBPG_GT_MFMA 1  /* int8 MFMA Gram tiles, exact int32 */
/* OR into the mode: keep the reference's slot-replacement rule
 * (CompGT.cpp:63-97) instead of the exact (dist, id) top-k.  That rule
 * decides which of several rows tied at the k-th distance survive, and its
 * stale-maximum step (CompGT.cpp:82-84) can evict a row of the true top-k;
 * this mode reproduces both. */
#define NBPG_GT_REFTIES 0x100
/* OR into the mode: inner-product ground truth (CompGT --use_mips,
 * CompGT.cpp:51, 147), any engine */
#define NBPG_GT_MIPS 0x200

/* ---- PCA: projection forms and code types.  The code type is what the
 *      reference derives from the output extension (PCACompress.cpp:25-37):
 *      .fbin = F32 (no codes), .u8bin, .i8bin (chunk 1), .u4bin, .i4bin
 *      (chunk 2, two codes per byte). ---- */
#define NBPG_PCA_PLAIN 0 /* fp32, products summed in index order: y bit-identical */
#define NBPG_PCA_BF16 1  /* bf16 MFMA, fp32 accumulation, quantizer fused */
#define NBPG_CODE_F32 0
#define NBPG_CODE_U8 1
#define NBPG_CODE_I8 2
#define NBPG_CODE_U4 3
#define NBPG_CODE_I4 4

/* (id, dist) record of the reference (query_params.h:97-104), 8 bytes */
typedef struct {
    uint32_t id;
    float dist;
} nbpg_iddist;

/* device-resident index: vectors + adjacency + per-row norms.
 * Replaces PointRange (point_range.h:33-325) and Graph (graph.h:119-400). */
typedef struct nbpg_index nbpg_index;

/* QueryParams (query_params.h:12-42), pure-quantized fields */
typedef struct {
    uint32_t k;            /* results per query; k <= L */
    double cut;            /* frontier cut factor, 1.35 in the sweep (check_nn_recall.h:213) */
    uint32_t L;            /* beamSize */
    uint64_t limit;        /* max expansions per query (|V| in the sweep) */
    uint32_t degree_limit; /* neighbours expanded per node (R in the sweep) */
} nbpg_query_params;

/* FAST-mode launch shape (the autotuner's knobs) */
typedef struct {
    uint32_t waves_per_group; /* queries per workgroup: 1..16 */
    uint32_t expand_width;    /* beam entries expanded per step: 1..4 */
    uint32_t filter_log2;     /* LDS visited-filter slots = 1 << filter_log2 (8..14) */
    uint32_t groups;          /* persistent workgroups; 0 = derive from occupancy (LDS, at
                                 most 24 waves per CU); always capped at what the device
                                 holds at once, as a wave's first query is its grid index */
    uint32_t row_split;       /* 0: one lane gathers a whole vector row; 1: one lane per
                                 16-byte chunk, a load instruction covers whole rows */
} nbpg_fast_config;

/* build_vamana's flags (Vamana.cpp:14-85, nsw.h:547-551) */
typedef struct {
    uint32_t R, L, VL, top_degree, num_split, num_passes;
    uint64_t limit; /* 0 = n */
    double min_alpha, max_alpha, min_cos_angle, max_cos_angle;
    int32_t dynamic_prune, rand_sp, repair_isolate;
    uint64_t sp_seed; /* seed replacing randomInt's random_device (utilities.h:55-63) */
} nbpg_build_params;

/* ---------------------------------------------------------------- */
const char *nbpg_last_error(void);
/* selects the device and caches its properties; call once per process */
int nbpg_init(int device);
int nbpg_device_info(int *cu_count, int *lds_per_group, int *l2_bytes);

/* ---- index lifecycle ----
 * stands in for load_point_range / PointRange ctor (point_range.h:17-57)
 * and Graph(max_degree, n) / Graph(file) (graph.h:163-221). */
int nbpg_index_create(nbpg_index **out, int dtype, uint64_t n, uint32_t dims, uint32_t R);
int nbpg_index_destroy(nbpg_index *idx);
/* the index's distance (NBPG_METRIC_L2 at creation): every search, prune,
 * build and distance call on it uses this metric afterwards */
int nbpg_index_set_metric(nbpg_index *idx, int metric);
/* the index's seed rows for NBPG_SEARCH_SEED_NEAREST: count (1..NBPG_MAX_SEEDS)
 * distinct host ids < n, kept in ascending order with a copy of their rows
 * taken now (count 0 drops them).  nbpg_index_set_points drops them too:
 * set them again after new rows.  No counterpart in the reference (its
 * callers pass start points per query through get_sp). */
int nbpg_index_set_seeds(nbpg_index *idx, const uint32_t *h_ids, uint32_t count);
/* upload rows [first, first+count) from host memory with the given stride
 * (the reference's 64-B padded rows, point_range.h:51-56, or packed) */
int nbpg_index_set_points(nbpg_index *idx, const void *host_rows, uint64_t host_stride,
                          uint64_t first, uint64_t count);
/* adjacency in the reference's in-memory layout: row i = [deg, id...],
 * R+1 u32 per row (graph.h:142-161); host memory */
int nbpg_index_set_graph(nbpg_index *idx, const uint32_t *host_rows, uint64_t first,
                         uint64_t count);
int nbpg_index_get_graph(const nbpg_index *idx, uint32_t *host_rows, uint64_t first,
                         uint64_t count);
/* Graph::save / Graph(file) (graph.h:169-221, 339-379) straight from / into
 * the resident adjacency: [u32 n][u32 R][u32 deg x n][ids].  load needs the
 * file's n to equal the index's and its R <= the index's R. */
int nbpg_index_save_graph(const nbpg_index *idx, const char *path);
int nbpg_index_load_graph(nbpg_index *idx, const char *path);
/* DFS::repair_isolate_node(G, sp) (dfs.h:101-253) on the resident graph:
 * nodes unreachable from sp get one reverse edge by the reference's four
 * rules, in ascending id order; *out_isolated = how many were unreachable. */
int nbpg_index_repair(nbpg_index *idx, uint32_t sp, uint64_t *out_isolated);
/* device views for callers that stage their own queries */
int nbpg_index_info(const nbpg_index *idx, int *dtype, uint64_t *n, uint32_t *dims, uint32_t *R,
                    uint32_t *row_stride);

/* ---- start points for the search's get_sp (check_nn_recall.h:282-297,
 * seeded into the beam as query_params.h:206-210 does): Lloyd k-means with
 * `count` centres (<= 256) over `sample` rows of the index drawn with
 * `seed`, `iters` steps, then the row nearest each centre.  Deterministic
 * (integer centre sums, FP64 distances, lowest index on ties).  Writes the
 * distinct ids, ascending, to host memory; *out_count <= count of them. */
int nbpg_pick_seeds(const nbpg_index *idx, uint32_t count, uint32_t sample, uint32_t iters,
                    uint64_t seed, uint32_t *h_out_ids, uint32_t *out_count);

/* ---- distances: Point::distance / euclidean_distance (point.h:35-120, 303-314)
 * out[i] = dist(query, row ids[i]) for i < n; d_query holds `dims` bytes. */
int nbpg_l2_batch(const nbpg_index *idx, const void *d_query, const uint32_t *d_ids, uint64_t n,
                  float *d_out, void *stream);
/* row-pair form over caller buffers: out[i] = dist(a_i, b_i), rows packed */
int nbpg_l2_pairs(int dtype, const void *d_a, const void *d_b, uint64_t n, uint32_t dims,
                  float *d_out, void *stream);

/* ---- search: qsearchAll (beamSearch.h:257-303) with get_sp = start points
 * (check_nn_recall.h:282-297).  d_queries: nq rows of `dims` bytes,
 * query_stride bytes apart.  Outputs nq*k ids / dists (id 0xFFFFFFFF past
 * the frontier), optional stats nq*2 = {visited, dist_cmps} as QueryStatic
 * (query_params.h:47-94). */
int nbpg_search(const nbpg_index *idx, const void *d_queries, uint64_t nq, uint64_t query_stride,
                const uint32_t *h_start_points, uint32_t sp_num, const nbpg_query_params *qp,
                int mode, const nbpg_fast_config *fast, uint32_t *d_out_ids, float *d_out_dists,
                uint32_t *d_out_stats, void *stream);
/* qsearchAll over quantized rows with full-precision rows beside them
 * (Base_Points != Q_Base_Points, beam_search_rerank, beamSearch.h:183-247).
 * The traversal uses the index (quantized) rows and d_queries; the raw rows
 * (raw_dtype, raw_dims elements each, packed n x raw_dims on the device)
 * and d_raw_queries (raw_query_stride bytes apart) give the final distances:
 *   rerank_factor > 0: rerank mode -- the first min(k * rerank_factor, fs)
 *     frontier entries are re-scored and the k nearest by (dist, id) kept;
 *   rerank_factor = 0: filtering mode -- every visited
 *     entry is re-scored and the first min(k, fs, visited) are kept; its
 *     dist_cmps counter adds those exact distances as the reference does.
 * NBPG_SEARCH_REF returns the reference's IDs (f32 rows: distances summed in
 * index order, point.h:137-146).  NBPG_SEARCH_FAST re-scores the top
 * min(k * rerank_factor, L) <= 1024 of its beam; with rerank_factor = 0
 * (filtering) it re-scores the query's visited list (NBPG_SEARCH_VISITED,
 * up to NBPG_VISITED_CAP(L) entries, L <= 256), keeps the first k by
 * (dist, id), and counts the re-scored entries in dist_cmps.  FAST uses the
 * index's workspace (calls on one index must not overlap).  Outputs nq x k. */
int nbpg_search_rerank(const nbpg_index *idx, const void *d_queries, uint64_t nq,
                       uint64_t query_stride, int raw_dtype, const void *d_raw_base,
                       uint32_t raw_dims, const void *d_raw_queries, uint64_t raw_query_stride,
                       const uint32_t *h_start_points, uint32_t sp_num,
                       const nbpg_query_params *qp, uint32_t rerank_factor, int mode,
                       const nbpg_fast_config *fast, uint32_t *d_out_ids, float *d_out_dists,
                       uint32_t *d_out_stats, void *stream);
/* ---- one nbpg_search captured as a HIP graph and replayed: the batched
 * driver qsearchAll (beamSearch.h:257-303) as a single graph launch per
 * batch.  create makes one eager call (argument checks), then captures
 * the same call; reference-mode state beyond LDS goes to scratch the graph
 * owns (never the index's workspace); launch replays it on `stream`, reading
 * whatever d_queries holds at that time and writing the same outputs.  The
 * index, buffers and start points are fixed at creation; the graph has its
 * own FAST work counter, which the kernel returns to zero, so the graph
 * holds no memset.  Launches of one graph must not overlap each other in
 * time (they share the counter and the outputs); other searches on the
 * index may run beside them. */
typedef struct nbpg_search_graph nbpg_search_graph;
int nbpg_search_graph_create(nbpg_search_graph **out, const nbpg_index *idx, const void *d_queries,
                             uint64_t nq, uint64_t query_stride, const uint32_t *h_start_points,
                             uint32_t sp_num, const nbpg_query_params *qp, int mode,
                             const nbpg_fast_config *fast, uint32_t *d_out_ids,
                             float *d_out_dists, uint32_t *d_out_stats);
int nbpg_search_graph_launch(nbpg_search_graph *g, void *stream);
int nbpg_search_graph_destroy(nbpg_search_graph *g);

/* nbpg_search_rerank without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, modes and results as nbpg_search_rerank, but
 * the scratch it needs -- FAST mode's nq x L candidate ids, reference mode's
 * per-workgroup search state when it does not fit in LDS (L >= 128 or so;
 * the grid then shrinks to what the scratch holds) -- lives in one fixed
 * array of NBPG_NOALLOC_SCRATCH_BYTES per GPU that the library holds until
 * the process ends, and no hipMalloc / hipFree or allocating library call is
 * made.  Limit: FAST mode needs nq * L * 4 <= NBPG_NOALLOC_SCRATCH_BYTES
 * (else NBPG_EINVAL); reference mode needs one query's state to fit (else
 * NBPG_EUNSUP).  Every call on a GPU shares that GPU's array: calls that
 * overlap in time on one GPU race (issue them on one stream). */
#define NBPG_NOALLOC_SCRATCH_BYTES (32ull << 20)
int nbpg_search_rerank_noalloc(const nbpg_index *idx, const void *d_queries, uint64_t nq,
                               uint64_t query_stride, int raw_dtype, const void *d_raw_base,
                               uint32_t raw_dims, const void *d_raw_queries,
                               uint64_t raw_query_stride, const uint32_t *h_start_points,
                               uint32_t sp_num, const nbpg_query_params *qp,
                               uint32_t rerank_factor, int mode, const nbpg_fast_config *fast,
                               uint32_t *d_out_ids, float *d_out_dists, uint32_t *d_out_stats,
                               void *stream);
/* FAST mode with per-query beam width and k (uneven batches); d_L, d_k are
 * nq u32 each on the device, outputs have stride out_k per query.  A query's
 * L is capped at max_L (and raised to 1), its k at min(L, out_k), by every
 * kernel shape alike.  Start points as nbpg_search's FAST mode (1..64
 * distinct host ids; more than one needs max_L <= 256). */
int nbpg_search_uneven(const nbpg_index *idx, const void *d_queries, uint64_t nq,
                       uint64_t query_stride, const uint32_t *d_L, const uint32_t *d_k,
                       uint32_t max_L, uint32_t out_k, const uint32_t *h_start_points,
                       uint32_t sp_num, const nbpg_fast_config *fast, uint32_t *d_out_ids,
                       float *d_out_dists, uint32_t *d_out_stats, void *stream);
/* nbpg_search_uneven captured as a HIP graph: the uneven batch of a
 * qsearchAll-style driver (beamSearch.h:257-303) as one graph launch.  Same
 * rules as nbpg_search_graph_*: create makes one eager call, then captures
 * it; a launch reads whatever d_queries, d_L and d_k hold at that time (the
 * per-query L and k may change between launches, capped by max_L and out_k
 * fixed at creation) and writes the same outputs; the graph owns its work
 * counter and holds no memset; launches of one graph must not overlap. */
int nbpg_search_uneven_graph_create(nbpg_search_graph **out, const nbpg_index *idx,
                                    const void *d_queries, uint64_t nq, uint64_t query_stride,
                                    const uint32_t *d_L, const uint32_t *d_k, uint32_t max_L,
                                    uint32_t out_k, const uint32_t *h_start_points,
                                    uint32_t sp_num, const nbpg_fast_config *fast,
                                    uint32_t *d_out_ids, float *d_out_dists,
                                    uint32_t *d_out_stats);

/* ---- RobustPrune: Prune::sort_candidates + Prune::robustPrune
 * (prune.h:49-87, 158-209) for a batch of nodes.  Node t's candidates are
 * d_cand[d_offsets[t] .. d_offsets[t+1]) with dist = d(node, cand); they are
 * deduped, sorted (dist, id), the node itself dropped, then pruned greedily
 * with alpha (d_alpha[t] if non-NULL, else alpha) and cos_angle.
 * Output: up to R ids/dists per node at stride R, count in d_out_counts. */
int nbpg_robust_prune_batch(const nbpg_index *idx, const uint32_t *d_nodes, uint64_t n_nodes,
                            const nbpg_iddist *d_cand, const uint64_t *d_offsets, uint32_t R,
                            const float *d_alpha, double alpha, double cos_angle,
                            uint32_t *d_out_ids, float *d_out_dists, uint32_t *d_out_counts,
                            void *stream);

/* ---- brute-force ground truth: compute_groundtruth (CompGT.cpp:29-134)
 * per query the k rows sorted by (dist, id), rows packed (dims bytes each).
 * mode = NBPG_GT_PLAIN or NBPG_GT_MFMA, optionally | NBPG_GT_REFTIES and
 * | NBPG_GT_MIPS. */
int nbpg_compgt(int dtype, const void *d_base, uint64_t n, const void *d_queries, uint64_t nq,
                uint32_t dims, uint32_t k, int mode, uint32_t *d_out_ids, float *d_out_dists,
                void *stream);

/* ---- rerank: the second half of beam_search_rerank (beamSearch.h:183-236).
 * d_cand holds each query's frontier from a search over quantized codes
 * (nq rows of cand_stride ids, first n_cand used, an id >= n ends a row).
 * The first min(k * rerank_factor, n_cand) candidates (at most 1024) are
 * re-scored against the full-precision rows (raw_dtype NBPG_F32 / NBPG_U8 /
 * NBPG_I8 / NBPG_U4X2 / NBPG_I4X2, dims elements each, packed) under
 * `metric` (NBPG_METRIC_L2 / NBPG_METRIC_MIPS) and the k nearest by
 * (dist, id) are written, nq x k; missing slots get NBPG_SENT / +inf.  f32
 * distances are the reference's sequential fp32 sums (point.h:137-156). */
int nbpg_rerank(int raw_dtype, int metric, const void *d_raw_base, uint64_t n, uint32_t dims,
                const void *d_raw_queries, uint64_t nq, const uint32_t *d_cand,
                uint32_t cand_stride, uint32_t n_cand, uint32_t k, uint32_t rerank_factor,
                uint32_t *d_out_ids, float *d_out_dists, void *stream);

/* ---- PCACompress (PCACompress.cpp:21-79): y = bias + W x per row
 * (matmul, matrix.h:86-97) and its codes (compress, matrix.h:99-127).
 * x: n rows of d_in elements of x_dtype (NBPG_F32, NBPG_U8 or NBPG_I8; the
 * byte types become floats as loadFloatPoint does, matrix.h:27-57).
 * W: d_out x d_in f32 row-major and bias: d_out f32, as loadMatrix reads a
 * .mat file (matrix.h:59-83; zeros when it has no bias).
 * code != NBPG_CODE_F32: d_out_codes gets n x (d_out / chunk) bytes;
 * d_out_y (optional there, required for NBPG_CODE_F32) gets y, n x d_out.
 * NBPG_PCA_BF16 needs d_in % 8 == 0, d_out a multiple of 16 up to 128. */
int nbpg_pca_project(int x_dtype, const void *d_x, uint64_t n, uint32_t d_in, const float *d_W,
                     const float *d_bias, uint32_t d_out, int mode, int code,
                     uint8_t *d_out_codes, float *d_out_y, void *stream);
/* compress (matrix.h:99-127) alone: y n x d f32 -> n x (d / chunk) codes */
int nbpg_pca_quantize(const float *d_y, uint64_t n, uint32_t d, int code, uint8_t *d_out,
                      void *stream);
/* FP64 statistics for a PCA fit (the reference's fit script,
 * py_data_tools/pca_process.py named at README.md:43, is not in its tree):
 * mean[d] and, if d_cov is non-NULL, the covariance d x d
 * sum (x - mean)(x - mean)^T / (n - 1), accumulated in FP64. */
int nbpg_pca_stats_f64(int x_dtype, const void *d_x, uint64_t n, uint32_t d, double *d_mean,
                       double *d_cov, void *stream);

/* ---- serving: a continuous-batching request loop over a resident index.
 * No counterpart in the reference (its searches run as whole batches,
 * check_nn_recall.h:74-76).  Requests of 1.. queries (host rows, own L and
 * k) are queued from any thread; a scheduler thread keeps `slots` batches in
 * flight on their own streams and admits queued requests, in arrival order,
 * into each slot that frees up, up to the slot's query capacity:
 * min(max_batch_queries, mem_budget / slots / bytes per query) where a query
 * takes row_stride + 8 + 8 * max_k device bytes (mem_budget 0 = no limit).
 * Each query runs the FAST search with its own L and k, and its result does
 * not depend on the other queries of its batch. ---- */
typedef struct {
    uint32_t slots;             /* batches in flight, 1..8 */
    uint32_t max_batch_queries; /* queries per batch */
    uint64_t mem_budget;        /* device bytes for all slots' buffers (0 = unbounded) */
    uint32_t max_L, max_k;      /* largest L and k a request may ask for */
    nbpg_fast_config fast;      /* launch shape of every batch */
    uint32_t sp_num;            /* start points: 1..64 distinct ids (more than one: max_L <= 256) */
    uint32_t start_points[64];  /* e.g. from nbpg_pick_seeds */
    uint32_t seed_nearest;      /* 1: each query starts from its nearest seed row of the index
                                   (NBPG_SEARCH_SEED_NEAREST; max_L <= 256), start points unused */
} nbpg_server_config;

typedef struct {
    uint64_t requests, queries, batches; /* completed since creation / last reset */
    double mean_batch_queries;
    double p50_ms, p99_ms, max_ms; /* request latency: submit -> results on the host */
    uint32_t batch_capacity;       /* queries per batch admitted */
} nbpg_server_stats_t;

typedef struct nbpg_server nbpg_server;
/* the index must outlive the server */
int nbpg_server_create(nbpg_server **out, const nbpg_index *idx, const nbpg_server_config *cfg);
/* copies the rows; returns a ticket to wait on (once) */
int nbpg_server_submit(nbpg_server *srv, const void *host_queries, uint32_t nq,
                       uint64_t query_stride, uint32_t L, uint32_t k, uint64_t *ticket);
/* blocks until the request is done; nq x k ids / dists (either may be NULL).
 * One wait per ticket: a second wait on a ticket that a thread is already
 * waiting on returns NBPG_EINVAL. */
int nbpg_server_wait(nbpg_server *srv, uint64_t ticket, uint32_t *host_ids, float *host_dists,
                     double *latency_ms);
/* a burst of n requests queued in one call (a front end that receives
 * several at once): request i is req_nq[i] consecutive rows of host_queries,
 * after request i-1's, all with the same L and k; tickets[i] as
 * nbpg_server_submit gives.  All n are checked before any is queued. */
int nbpg_server_submit_many(nbpg_server *srv, const void *host_queries, uint32_t n,
                            const uint32_t *req_nq, uint64_t query_stride, uint32_t L, uint32_t k,
                            uint64_t *tickets);
/* nbpg_server_wait for n tickets in turn: request i's nq x k ids / dists
 * follow request i-1's in host_ids / host_dists (either may be NULL),
 * latency_ms[i] its latency (NULL allowed).  A failed request keeps its
 * place and the first failure's status is returned after all n; an unknown
 * ticket stops the call there (the later tickets stay to be waited on). */
int nbpg_server_wait_many(nbpg_server *srv, uint32_t n, const uint64_t *tickets,
                          uint32_t *host_ids, float *host_dists, double *latency_ms);
int nbpg_server_stats(nbpg_server *srv, int reset, nbpg_server_stats_t *stats);
/* finishes queued requests, lets threads already blocked in
 * nbpg_server_wait take their results, then frees the slots and the results
 * nobody waited for.  No submit or wait may START once destroy has been
 * called. */
int nbpg_server_destroy(nbpg_server *srv);

/* ---- graph build: build_vamana (nsw.h:547-663) on the device; the
 * index's points must be set; the graph is (re)built in place (set a graph
 * first to continue from it, Vamana.cpp:174-177). */
int nbpg_build(nbpg_index *idx, const nbpg_build_params *bp, uint64_t *out_isolated);

#ifdef __cplusplus
}
#endif
#endif
