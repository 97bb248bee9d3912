// ref_shim.cpp -- extern "C" entry points over the reference's own headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// sources where they lie (include path into the reference tree); the output
// goes to oracle/_ref/ and is used only to pin the C restatement in
// oracle/nbpg_oracle.c and to write the golden fixtures under tests/golden/.
// No reference text is copied here: every call below goes into the
// reference's templates (namespace ant).
#include <algorithm>
#include <cstdint>
#include <functional>
#include <iostream>
#include <string>
#include <cstring>
#include <random>
#include <chrono>
#include <omp.h>
#include <vector>

#include "utils/Point/point.h"
#include "utils/Point/point_range.h"
#include "utils/Graph/graph.h"
#include "utils/Graph/query_params.h"
#include "utils/Graph/beamSearch.h"
#include "utils/Graph/prune.h"
#include "utils/Graph/dfs.h"
#include "utils/utilities.h"
#include "data_proc/base/matrix.h"

using namespace ant;

namespace {

// the metric of the current call: a dtype argument | 0x100 selects the
// inner product (DistanceMetric::MIPS) for every range the call builds
DistanceMetric g_metric = DistanceMetric::EUCLIDEAN;

int take_metric(int dtype)
{
    g_metric = (dtype & 0x100) ? DistanceMetric::MIPS : DistanceMetric::EUCLIDEAN;
    return dtype & 0xFF;
}

template <typename T>
PointRange<Point<T>> make_points(const uint8_t *data, uint64_t n, uint32_t dims)
{
    using PR = PointRange<Point<T>>;
    PR P(n, typename PR::Parameters(dims, g_metric));
    const size_t nb = P.params.num_bytes();
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(reinterpret_cast<uint8_t *>(P[i].data()), data + i * nb, nb);
    return P;
}

Graph<uint32_t> make_graph(const uint32_t *rows, uint64_t n, uint32_t R)
{
    Graph<uint32_t> G(R, n, false);
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(G[i].data(), rows + i * (R + 1), sizeof(uint32_t) * (R + 1));
    return G;
}

template <typename T>
float dist_t(const void *a, const void *b, uint32_t dims)
{
    if (g_metric == DistanceMetric::MIPS)
        return mips_distance(reinterpret_cast<const T *>(a), reinterpret_cast<const T *>(b), dims);
    return euclidean_distance(reinterpret_cast<const T *>(a), reinterpret_cast<const T *>(b), dims);
}

template <typename T>
int search_t(const uint8_t *base, uint64_t n, uint32_t dims, const uint32_t *rows, uint32_t R,
             const uint8_t *queries, uint64_t nq, uint32_t L, uint32_t k, double cut,
             uint64_t limit, uint32_t degree_limit, uint32_t entry, uint32_t *out_ids,
             uint32_t *out_stats)
{
    auto B = make_points<T>(base, n, dims);
    auto Q = make_points<T>(queries, nq, dims);
    auto G = make_graph(rows, n, R);
    QueryStatic stats(nq);
    stats.clear();
    uint32_t sp = entry;
    std::function<std::pair<uint32_t *, uint32_t>(uint32_t)> get_sp =
        [&](uint32_t) { return std::make_pair(&sp, (uint32_t)1); };
    QueryParams QP(k, L, cut, limit, degree_limit);
    auto res = qsearchAll<uint32_t>(Q, Q, G, B, B, stats, get_sp, QP);
    for (uint64_t i = 0; i < nq; ++i) {
        for (uint32_t j = 0; j < k; ++j)
            out_ids[i * k + j] = j < res[i].size() ? res[i][j] : 0xFFFFFFFFu;
        out_stats[2 * i] = stats.visited[i];
        out_stats[2 * i + 1] = stats.distances[i];
    }
    return 0;
}

// qsearchAll with the raw ranges beside the quantized ones
// (beam_search_rerank picks rerank / filtering from rerank_factor)
template <typename T, typename X>
int search_exact_t(const uint8_t *base, uint64_t n, uint32_t dims, const uint32_t *rows,
                   uint32_t R, const uint8_t *queries, uint64_t nq, const uint8_t *raw_base,
                   uint32_t raw_dims, const uint8_t *raw_queries, uint32_t L, uint32_t k,
                   double cut, uint64_t limit, uint32_t degree_limit, uint32_t entry,
                   uint32_t rerank_factor, uint32_t *out_ids, uint32_t *out_stats)
{
    auto B = make_points<T>(base, n, dims);
    auto Q = make_points<T>(queries, nq, dims);
    auto XB = make_points<X>(raw_base, n, raw_dims);
    auto XQ = make_points<X>(raw_queries, nq, raw_dims);
    auto G = make_graph(rows, n, R);
    QueryStatic stats(nq);
    stats.clear();
    uint32_t sp = entry;
    std::function<std::pair<uint32_t *, uint32_t>(uint32_t)> get_sp =
        [&](uint32_t) { return std::make_pair(&sp, (uint32_t)1); };
    QueryParams QP(k, L, cut, limit, degree_limit, rerank_factor);
    auto res = qsearchAll<uint32_t>(XQ, Q, G, XB, B, stats, get_sp, QP);
    for (uint64_t i = 0; i < nq; ++i) {
        for (uint32_t j = 0; j < k; ++j)
            out_ids[i * k + j] = j < res[i].size() ? res[i][j] : 0xFFFFFFFFu;
        out_stats[2 * i] = stats.visited[i];
        out_stats[2 * i + 1] = stats.distances[i];
    }
    return 0;
}

template <typename T>
int search_exact_x(int raw_dtype, const uint8_t *base, uint64_t n, uint32_t dims,
                   const uint32_t *rows, uint32_t R, const uint8_t *queries, uint64_t nq,
                   const uint8_t *raw_base, uint32_t raw_dims, const uint8_t *raw_queries,
                   uint32_t L, uint32_t k, double cut, uint64_t limit, uint32_t degree_limit,
                   uint32_t entry, uint32_t rerank_factor, uint32_t *out_ids, uint32_t *out_stats)
{
#define NBPGR_X(XT)                                                                                \
    search_exact_t<T, XT>(base, n, dims, rows, R, queries, nq, raw_base, raw_dims, raw_queries, L, \
                          k, cut, limit, degree_limit, entry, rerank_factor, out_ids, out_stats)
    switch (raw_dtype) {
    case 0: return NBPGR_X(uint8_t);
    case 1: return NBPGR_X(int8_t);
    case 4: return NBPGR_X(float);
    }
#undef NBPGR_X
    return -1;
}

template <typename T>
uint64_t prune_t(const uint8_t *base, uint64_t n, uint32_t dims, uint32_t p, IdDist<uint32_t> *cand,
                 uint64_t size, uint32_t R, double alpha, double cos_angle)
{
    auto B = make_points<T>(base, n, dims);
    Graph<uint32_t> G(1, 1, false);
    Prune pr(R, 1);
    size_t sz = size;
    pr.sort_candidates(p, G, cand, sz, 0.0f);
    pr.robustPrune(p, cand, sz, B, alpha, cos_angle, 1);
    return sz;
}

// many nodes over one PointRange; OpenMP over nodes like the reference's
// build loops; returns the seconds of the prune loop alone
template <typename T>
double prune_batch_t(const uint8_t *base, uint64_t n, uint32_t dims, const uint32_t *nodes,
                     uint64_t P, IdDist<uint32_t> *cand, uint32_t C, uint32_t R, double alpha,
                     double cos_angle, uint32_t *out_counts)
{
    auto B = make_points<T>(base, n, dims);
    Graph<uint32_t> G(1, 1, false);
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel
    {
        Prune pr(R, 1);
#pragma omp for schedule(dynamic, 64)
        for (int64_t t = 0; t < (int64_t)P; ++t) {
            size_t sz = C;
            IdDist<uint32_t> *c = cand + (size_t)t * C;
            pr.sort_candidates(nodes[t], G, c, sz, 0.0f);
            pr.robustPrune(nodes[t], c, sz, B, alpha, cos_angle, 1);
            out_counts[t] = (uint32_t)sz;
        }
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

extern "C" {

float nbpgr_distance(int dtype, const void *a, const void *b, uint32_t dims)
{
    dtype = take_metric(dtype);
    switch (dtype) {
    case 0: return dist_t<uint8_t>(a, b, dims);
    case 1: return dist_t<int8_t>(a, b, dims);
    case 2: return dist_t<uint4_2>(a, b, dims);
    case 3: return dist_t<int4_2>(a, b, dims);
    case 4: return dist_t<float>(a, b, dims);
    }
    return 0.0f;
}

uint64_t nbpgr_hash64_2(uint64_t x) { return hash64_2(x); }

uint32_t nbpgr_hash_filter_size(uint32_t beam)
{
    return (uint32_t)BeamSearchMemoryCell<uint32_t>::get_max_hash_filter_size(beam);
}

int nbpgr_search(int dtype, const uint8_t *base, uint64_t n, uint32_t dims, const uint32_t *rows,
                 uint32_t R, const uint8_t *queries, uint64_t nq, uint32_t L, uint32_t k,
                 double cut, uint64_t limit, uint32_t degree_limit, uint32_t entry,
                 uint32_t *out_ids, uint32_t *out_stats)
{
    dtype = take_metric(dtype);
    switch (dtype) {
    case 0: return search_t<uint8_t>(base, n, dims, rows, R, queries, nq, L, k, cut, limit, degree_limit, entry, out_ids, out_stats);
    case 1: return search_t<int8_t>(base, n, dims, rows, R, queries, nq, L, k, cut, limit, degree_limit, entry, out_ids, out_stats);
    case 2: return search_t<uint4_2>(base, n, dims, rows, R, queries, nq, L, k, cut, limit, degree_limit, entry, out_ids, out_stats);
    case 3: return search_t<int4_2>(base, n, dims, rows, R, queries, nq, L, k, cut, limit, degree_limit, entry, out_ids, out_stats);
    }
    return -1;
}

int nbpgr_search_exact(int dtype, const uint8_t *base, uint64_t n, uint32_t dims,
                       const uint32_t *rows, uint32_t R, const uint8_t *queries, uint64_t nq,
                       int raw_dtype, const uint8_t *raw_base, uint32_t raw_dims,
                       const uint8_t *raw_queries, uint32_t L, uint32_t k, double cut,
                       uint64_t limit, uint32_t degree_limit, uint32_t entry,
                       uint32_t rerank_factor, uint32_t *out_ids, uint32_t *out_stats)
{
    dtype = take_metric(dtype);
    switch (dtype) {
    case 0: return search_exact_x<uint8_t>(raw_dtype, base, n, dims, rows, R, queries, nq, raw_base, raw_dims, raw_queries, L, k, cut, limit, degree_limit, entry, rerank_factor, out_ids, out_stats);
    case 2: return search_exact_x<uint4_2>(raw_dtype, base, n, dims, rows, R, queries, nq, raw_base, raw_dims, raw_queries, L, k, cut, limit, degree_limit, entry, rerank_factor, out_ids, out_stats);
    }
    return -1;
}

// DFS::repair_isolate_node(G, sp) on reference-layout rows, in place
uint64_t nbpgr_repair(uint32_t *rows, uint64_t n, uint32_t R, uint32_t sp)
{
    auto G = make_graph(rows, n, R);
    const size_t cnt = DFS::repair_isolate_node(G, (size_t)sp);
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(rows + i * (R + 1), G[i].data(), sizeof(uint32_t) * (R + 1));
    return cnt;
}

uint64_t nbpgr_prune(int dtype, const uint8_t *base, uint64_t n, uint32_t dims, uint32_t p,
                     void *cand, uint64_t size, uint32_t R, double alpha, double cos_angle)
{
    dtype = take_metric(dtype);
    auto *c = reinterpret_cast<IdDist<uint32_t> *>(cand);
    switch (dtype) {
    case 0: return prune_t<uint8_t>(base, n, dims, p, c, size, R, alpha, cos_angle);
    case 1: return prune_t<int8_t>(base, n, dims, p, c, size, R, alpha, cos_angle);
    case 2: return prune_t<uint4_2>(base, n, dims, p, c, size, R, alpha, cos_angle);
    case 3: return prune_t<int4_2>(base, n, dims, p, c, size, R, alpha, cos_angle);
    }
    return 0;
}

double nbpgr_prune_batch(int dtype, const uint8_t *base, uint64_t n, uint32_t dims,
                         const uint32_t *nodes, uint64_t P, void *cand, uint32_t C, uint32_t R,
                         double alpha, double cos_angle, uint32_t *out_counts)
{
    dtype = take_metric(dtype);
    auto *c = reinterpret_cast<IdDist<uint32_t> *>(cand);
    switch (dtype) {
    case 0: return prune_batch_t<uint8_t>(base, n, dims, nodes, P, c, C, R, alpha, cos_angle, out_counts);
    case 1: return prune_batch_t<int8_t>(base, n, dims, nodes, P, c, C, R, alpha, cos_angle, out_counts);
    case 2: return prune_batch_t<uint4_2>(base, n, dims, nodes, P, c, C, R, alpha, cos_angle, out_counts);
    case 3: return prune_batch_t<int4_2>(base, n, dims, nodes, P, c, C, R, alpha, cos_angle, out_counts);
    }
    return -1.0;
}

void nbpgr_shuffle(uint32_t *a, uint64_t n, uint32_t seed)
{
    std::vector<uint32_t> v(a, a + n);
    std::shuffle(v.begin(), v.end(), std::mt19937(seed));
    std::memcpy(a, v.data(), sizeof(uint32_t) * n);
}

void nbpgr_matmul(const float *x, uint64_t n, uint32_t d_in, const float *W, const float *bias,
                  uint32_t d_out, float *y)
{
    using PR = PointRange<Point<float>>;
    PR from(n, PR::Parameters(d_in, DistanceMetric::MIPS));
    PR to(n, PR::Parameters(d_out));
    PR mat(d_out, PR::Parameters(d_in, DistanceMetric::MIPS));
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(from[i].data(), x + i * d_in, sizeof(float) * d_in);
    for (uint32_t o = 0; o < d_out; ++o)
        std::memcpy(mat[o].data(), W + (size_t)o * d_in, sizeof(float) * d_in);
    std::vector<float> b(bias, bias + d_out);
    matmul(from, to, mat, b);
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(y + i * d_out, to[i].data(), sizeof(float) * d_out);
}

void nbpgr_compress(const float *y, uint64_t n, uint32_t d, int is_signed, uint32_t chunk,
                    uint8_t *out)
{
    using PR = PointRange<Point<float>>;
    using PU = PointRange<Point<uint8_t>>;
    PR to(n, PR::Parameters(d));
    const uint32_t od = d / chunk;
    PU cp(n, PU::Parameters(od));
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(to[i].data(), y + i * d, sizeof(float) * d);
    compress(to, cp, is_signed ? "int" : "uint", chunk);
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(out + i * od, cp[i].data(), od);
}

// persistent u8 index for timing: the copy into the reference's containers
// happens once in nbpgr_open; nbpgr_timed_search brackets qsearchAll alone,
// as checkRecall does (check_nn_recall.h:74-76)
struct RefHandle {
    PointRange<Point<uint8_t>> B;
    Graph<uint32_t> G;
    uint32_t dims;
};

void *nbpgr_open(const uint8_t *base, uint64_t n, uint32_t dims, const uint32_t *rows, uint32_t R)
{
    auto *h = new RefHandle{make_points<uint8_t>(base, n, dims), make_graph(rows, n, R), dims};
    return h;
}

double nbpgr_timed_search(void *hv, const uint8_t *queries, uint64_t nq, uint32_t L, uint32_t k,
                          double cut, uint64_t limit, uint32_t degree_limit, uint32_t entry,
                          uint32_t *out_ids, uint32_t *out_stats)
{
    auto *h = reinterpret_cast<RefHandle *>(hv);
    auto Q = make_points<uint8_t>(queries, nq, h->dims);
    QueryStatic stats(nq);
    stats.clear();
    uint32_t sp = entry;
    std::function<std::pair<uint32_t *, uint32_t>(uint32_t)> get_sp =
        [&](uint32_t) { return std::make_pair(&sp, (uint32_t)1); };
    QueryParams QP(k, L, cut, limit, degree_limit);
    auto t0 = std::chrono::steady_clock::now();
    auto res = qsearchAll<uint32_t>(Q, Q, h->G, h->B, h->B, stats, get_sp, QP);
    auto t1 = std::chrono::steady_clock::now();
    for (uint64_t i = 0; i < nq; ++i) {
        for (uint32_t j = 0; j < k; ++j)
            out_ids[i * k + j] = j < res[i].size() ? res[i][j] : 0xFFFFFFFFu;
        if (out_stats) {
            out_stats[2 * i] = stats.visited[i];
            out_stats[2 * i + 1] = stats.distances[i];
        }
    }
    return std::chrono::duration<double>(t1 - t0).count();
}

void nbpgr_close(void *hv) { delete reinterpret_cast<RefHandle *>(hv); }

int nbpgr_threads(void) { return omp_get_max_threads(); }

} // extern "C"
