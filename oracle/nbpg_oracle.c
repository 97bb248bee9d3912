/*
 * nbpg_oracle.c -- CPU restatement of the reference NBPG hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).
 * Each function follows one reference routine; citations are file:line in
 * the reference tree.  Written in plain C11 + OpenMP from the reference's
 * behaviour.  Deliberate divergence: the build's random start points
 * (fillMergeSP, nsw.h:584-599) come from a seeded splitmix64 stream instead
 * of std::random_device-seeded std::mt19937 (utilities.h:55-63), so builds
 * are reproducible.
 */
#include "nbpg_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* distances: point.h:27-156                                           */
/* ------------------------------------------------------------------ */

static float l2_u8(const uint8_t *p, const uint8_t *q, uint32_t d)
{
    int32_t s = 0;
    for (uint32_t i = 0; i < d; ++i) {
        int32_t t = (int32_t)p[i] - (int32_t)q[i];
        s += t * t;
    }
    return (float)s;
}

static float l2_i8(const int8_t *p, const int8_t *q, uint32_t d)
{
    int32_t s = 0;
    for (uint32_t i = 0; i < d; ++i) {
        int32_t t = (int32_t)p[i] - (int32_t)q[i];
        s += t * t;
    }
    return (float)s;
}

/* point.h:27-32, 77-90, 107-120: both packed types split a byte into
 * hi = b & 0xF0 and lo = (uint8)(b << 4), read as unsigned */
static float l2_4bit(const uint8_t *p, const uint8_t *q, uint32_t d)
{
    int32_t s = 0;
    for (uint32_t i = 0; i < d; ++i) {
        int32_t ph = p[i] & 0xF0, pl = (uint8_t)(p[i] << 4);
        int32_t qh = q[i] & 0xF0, ql = (uint8_t)(q[i] << 4);
        s += (ph - qh) * (ph - qh);
        s += (pl - ql) * (pl - ql);
    }
    return (float)s;
}

static float l2_f32(const float *p, const float *q, uint32_t d)
{
    float s = 0.0f;
    for (uint32_t i = 0; i < d; ++i) {
        float t = p[i] - q[i];
        s += t * t;
    }
    return s;
}

/* mips_distance (point.h:46-156): the negated inner product, integer types
 * summed exactly in int32, f32 in index order (nibbles as above) */
static float ip_int(int dtype, const uint8_t *p, const uint8_t *q, uint32_t d)
{
    int32_t s = 0;
    for (uint32_t i = 0; i < d; ++i) {
        if (dtype == NBPGO_U8)
            s += (int32_t)p[i] * (int32_t)q[i];
        else if (dtype == NBPGO_I8)
            s += (int32_t)(int8_t)p[i] * (int32_t)(int8_t)q[i];
        else {
            s += (int32_t)(p[i] & 0xF0) * (int32_t)(q[i] & 0xF0);
            s += (int32_t)(uint8_t)(p[i] << 4) * (int32_t)(uint8_t)(q[i] << 4);
        }
    }
    return -(float)s;
}

static float ip_f32(const float *p, const float *q, uint32_t d)
{
    float s = 0.0f;
    for (uint32_t i = 0; i < d; ++i)
        s += p[i] * q[i];
    return -s;
}

float nbpgo_distance(int dtype, const void *a, const void *b, uint32_t dims)
{
    if (dtype & NBPGO_MIPS) {
        dtype &= ~NBPGO_MIPS;
        if (dtype == NBPGO_F32)
            return ip_f32((const float *)a, (const float *)b, dims);
        return ip_int(dtype, (const uint8_t *)a, (const uint8_t *)b, dims);
    }
    switch (dtype) {
    case NBPGO_U8: return l2_u8((const uint8_t *)a, (const uint8_t *)b, dims);
    case NBPGO_I8: return l2_i8((const int8_t *)a, (const int8_t *)b, dims);
    case NBPGO_U4X2:
    case NBPGO_I4X2: return l2_4bit((const uint8_t *)a, (const uint8_t *)b, dims);
    case NBPGO_F32: return l2_f32((const float *)a, (const float *)b, dims);
    default: return 0.0f;
    }
}

static inline const uint8_t *row_of(const nbpgo_points *P, uint64_t i)
{
    return P->data + i * P->stride;
}

static inline float pdist(const nbpgo_points *P, uint64_t a, const uint8_t *q)
{
    return nbpgo_distance(P->dtype, row_of(P, a), q, P->dims);
}

/* ------------------------------------------------------------------ */
/* hashing: utilities.h:31-37; filter size query_params.h:243-247      */
/* ------------------------------------------------------------------ */

uint64_t nbpgo_hash64_2(uint64_t x)
{
    x = (x ^ (x >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    x = (x ^ (x >> 27)) * UINT64_C(0x94d049bb133111eb);
    return x ^ (x >> 31);
}

uint32_t nbpgo_hash_filter_size(uint32_t beam)
{
    uint32_t sq = beam * beam; /* uint32 product, as in the reference */
    int bits = (int)(ceil(log2((double)sq)) - 2.0);
    if (bits < 10)
        bits = 10;
    return 1u << bits;
}

/* (dist, id) order: query_params.h:115-119 */
static inline int less_id(const nbpgo_iddist *a, const nbpgo_iddist *b)
{
    return a->dist < b->dist || (a->dist == b->dist && a->id < b->id);
}

static int cmp_iddist(const void *x, const void *y)
{
    const nbpgo_iddist *a = (const nbpgo_iddist *)x, *b = (const nbpgo_iddist *)y;
    if (less_id(a, b))
        return -1;
    if (less_id(b, a))
        return 1;
    return 0;
}

static int cmp_by_id(const void *x, const void *y)
{
    const nbpgo_iddist *a = (const nbpgo_iddist *)x, *b = (const nbpgo_iddist *)y;
    if (a->id != b->id)
        return a->id < b->id ? -1 : 1;
    /* equal ids carry equal distances on every path that reaches here */
    return cmp_iddist(x, y);
}

/* ------------------------------------------------------------------ */
/* CompGT: CompGT.cpp:29-134                                            */
/* ------------------------------------------------------------------ */

static size_t elem_size(int dtype) { return (dtype & ~NBPGO_MIPS) == NBPGO_F32 ? 4 : 1; }

int nbpgo_compgt(int dtype, const void *base, uint64_t n, const void *queries, uint64_t nq,
                 uint32_t dims, uint32_t k, int tie_mode, uint32_t *out_ids, float *out_dists,
                 int nthreads)
{
    if (k == 0 || n < k)
        return -1;
    const size_t rb = (size_t)dims * elem_size(dtype);
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    int err = 0;
#pragma omp parallel
    {
        nbpgo_iddist *top = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * k);
        if (!top) {
#pragma omp atomic write
            err = -2;
        }
#pragma omp for schedule(dynamic, 4)
        for (int64_t j = 0; j < (int64_t)nq; ++j) {
            if (!top)
                continue;
            const uint8_t *q = (const uint8_t *)queries + (size_t)j * rb;
            uint32_t size = 0;
            if (tie_mode == 0) {
                /* CompGT.cpp:67-97: fill k slots tracking the first max
                 * slot (strict >), then admit dist < max into that slot
                 * and rescan for the first max slot */
                float topdist = -FLT_MAX;
                uint32_t toppos = 0;
                for (uint64_t i = 0; i < n; ++i) {
                    float d = nbpgo_distance(dtype, (const uint8_t *)base + i * rb, q, dims);
                    if (size < k) {
                        if (d > topdist) {
                            topdist = d;
                            toppos = size;
                        }
                        top[size].id = (uint32_t)i;
                        top[size].dist = d;
                        ++size;
                    } else if (d < topdist) {
                        /* CompGT.cpp:82-84 reads slot 0 before the
                         * eviction writes: when the max sat in slot 0 the
                         * rescan starts from the evicted value */
                        float nd = top[0].dist;
                        uint32_t np = 0;
                        top[toppos].id = (uint32_t)i;
                        top[toppos].dist = d;
                        for (uint32_t l = 1; l < size; ++l)
                            if (top[l].dist > nd) {
                                nd = top[l].dist;
                                np = l;
                            }
                        topdist = nd;
                        toppos = np;
                    }
                }
            } else {
                /* exact (dist, id) top-k: a sorted insertion list */
                for (uint64_t i = 0; i < n; ++i) {
                    nbpgo_iddist c = {(uint32_t)i,
                                      nbpgo_distance(dtype, (const uint8_t *)base + i * rb, q, dims)};
                    if (size == k && !less_id(&c, &top[k - 1]))
                        continue;
                    uint32_t pos = size < k ? size : k - 1;
                    while (pos > 0 && less_id(&c, &top[pos - 1])) {
                        top[pos] = top[pos - 1];
                        --pos;
                    }
                    top[pos] = c;
                    if (size < k)
                        ++size;
                }
            }
            /* CompGT.cpp:107-123: sort (dist, id), emit k */
            qsort(top, size, sizeof(nbpgo_iddist), cmp_iddist);
            for (uint32_t l = 0; l < k; ++l) {
                out_ids[(size_t)j * k + l] = top[l].id;
                out_dists[(size_t)j * k + l] = top[l].dist;
            }
        }
        free(top);
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* Prune: prune.h:49-74 (sort_candidates), 158-209 (limitedSimilarity) */
/* ------------------------------------------------------------------ */

static uint64_t dedupe_by_index(nbpgo_iddist *c, uint64_t n)
{
    /* prune.h:251-260: sort by id, keep the first of each id run */
    if (n == 0)
        return 0;
    qsort(c, n, sizeof(nbpgo_iddist), cmp_by_id);
    uint64_t w = 1;
    for (uint64_t r = 1; r < n; ++r)
        if (c[r].id != c[w - 1].id)
            c[w++] = c[r];
    return w;
}

static void sort_with_p(nbpgo_iddist *c, uint64_t n, uint32_t p)
{
    /* biased_less with dir_bias 0 (prune.h:55-69): p sorts as +FLT_MAX.
     * Insertion sort keeps this reentrant; candidate lists are short. */
    for (uint64_t i = 1; i < n; ++i) {
        nbpgo_iddist v = c[i];
        nbpgo_iddist vk = v;
        if (vk.id == p)
            vk.dist = FLT_MAX;
        uint64_t j = i;
        while (j > 0) {
            nbpgo_iddist u = c[j - 1];
            if (u.id == p)
                u.dist = FLT_MAX;
            if (!less_id(&vk, &u))
                break;
            c[j] = c[j - 1];
            --j;
        }
        c[j] = v;
    }
}

uint64_t nbpgo_sort_candidates(uint32_t p, nbpgo_iddist *cand, uint64_t n)
{
    n = dedupe_by_index(cand, n);
    sort_with_p(cand, n, p);
    while (n > 0 && cand[n - 1].id == p) /* prune.h:72-73 */
        --n;
    return n;
}

#define KEPS 1e-5

static float cosine_at_p(float d_ps, float d_pp, float d_sp)
{
    /* prune.h:218-225 */
    float b = sqrtf(d_pp), c = sqrtf(d_ps);
    if (b == 0.f || c == 0.f)
        return 1.f;
    return (d_pp + d_ps - d_sp) / (2.f * b * c);
}

uint64_t nbpgo_robust_prune(uint32_t p, nbpgo_iddist *cand, uint64_t *n, const nbpgo_points *P,
                            uint32_t R, double alpha, double cos_angle, uint64_t start_prune_id)
{
    uint64_t comps = 0;
    if (alpha <= KEPS && cos_angle <= KEPS) { /* prune.h:164-169 */
        if (*n > R)
            *n = R;
        return 0;
    }
    uint64_t sel = 0, scan = 0, size = *n;
    while (sel < R && scan < size) {
        uint32_t ps = cand[scan].id;
        ++scan;
        if (ps == p || ps == 0xFFFFFFFFu)
            continue;
        cand[sel++] = cand[scan - 1];
        const float d_ps = cand[scan - 1].dist;
        for (uint64_t i = scan; i < size; ++i) {
            if (i < start_prune_id)
                continue;
            uint32_t pp = cand[i].id;
            if (pp == 0xFFFFFFFFu)
                continue;
            ++comps;
            float d_sp = nbpgo_distance(P->dtype, row_of(P, ps), row_of(P, pp), P->dims);
            float d_pp = cand[i].dist;
            /* prune.h:227-238 */
            if (alpha > KEPS && alpha * (double)d_sp <= (double)d_pp)
                cand[i].id = 0xFFFFFFFFu;
            else if (cos_angle > KEPS && cosine_at_p(d_ps, d_pp, d_sp) > cos_angle)
                cand[i].id = 0xFFFFFFFFu;
        }
    }
    *n = sel;
    return comps;
}

/* ------------------------------------------------------------------ */
/* Beam search: query_params.h:131-260, beamSearch.h:43-171             */
/* ------------------------------------------------------------------ */

typedef struct {
    uint32_t L;
    nbpgo_iddist *frontier;
    uint64_t frontier_size;
    nbpgo_iddist *new_frontier; /* 2L + R */
    nbpgo_iddist *unvisited;    /* L */
    uint32_t *hash;
    uint64_t hash_size;
    uint32_t *filtered;       /* R */
    nbpgo_iddist *candidates; /* L + R */
    nbpgo_iddist *visited;
    uint64_t visited_size, max_visited;
    nbpgo_iddist *copy_visited; /* max_visited (+R slack) */
    float *tmp_visited;         /* filtering mode: exact distance of visited[i] */
    uint64_t num_visited;
} cell_t;

static int cell_alloc(cell_t *c, uint32_t L, uint32_t R, uint64_t max_visited)
{
    memset(c, 0, sizeof(*c));
    c->L = L;
    c->max_visited = max_visited;
    c->hash_size = nbpgo_hash_filter_size(L);
    c->frontier = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * (L + R + 1));
    c->new_frontier = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * (2 * (size_t)L + R + 1));
    c->unvisited = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * (L + 1));
    c->hash = (uint32_t *)malloc(sizeof(uint32_t) * c->hash_size);
    c->filtered = (uint32_t *)malloc(sizeof(uint32_t) * (R + 1));
    c->candidates = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * ((size_t)L + R + 1));
    c->visited = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * (max_visited + 1));
    c->copy_visited = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * (max_visited + R + 1));
    c->tmp_visited = (float *)malloc(sizeof(float) * (max_visited + 1));
    return (c->frontier && c->new_frontier && c->unvisited && c->hash && c->filtered &&
            c->candidates && c->visited && c->copy_visited && c->tmp_visited)
               ? 0
               : -2;
}

static void cell_free(cell_t *c)
{
    free(c->frontier);
    free(c->new_frontier);
    free(c->unvisited);
    free(c->hash);
    free(c->filtered);
    free(c->candidates);
    free(c->visited);
    free(c->copy_visited);
    free(c->tmp_visited);
}

static void cell_clear(cell_t *c)
{
    /* query_params.h:173-180 */
    c->frontier_size = 0;
    c->visited_size = 0;
    c->num_visited = 0;
    memset(c->hash, 0xFF, sizeof(uint32_t) * c->hash_size);
}

/* query_params.h:223-230: lossy direct-mapped filter, overwrite on miss */
static inline int has_been_seen(cell_t *c, uint32_t a)
{
    uint64_t loc = nbpgo_hash64_2((uint64_t)a) & (c->hash_size - 1);
    if (c->hash[loc] == a)
        return 1;
    c->hash[loc] = a;
    return 0;
}

/* query_params.h:203-221 */
static void init_starting_points(cell_t *c, const nbpgo_points *P, const uint8_t *q,
                                 const uint32_t *sp, uint32_t sp_num)
{
    for (uint32_t s = 0; s < sp_num; ++s) {
        c->frontier[s].id = sp[s];
        c->frontier[s].dist = pdist(P, sp[s], q);
        has_been_seen(c, sp[s]);
    }
    c->frontier_size = sp_num;
    qsort(c->frontier, sp_num, sizeof(nbpgo_iddist), cmp_iddist);
    memcpy(c->unvisited, c->frontier, sizeof(nbpgo_iddist) * sp_num);
}

static inline uint32_t g_deg(const nbpgo_graph *G, uint64_t v) { return G->rows[v * (G->R + 1)]; }
static inline uint32_t *g_nbrs(const nbpgo_graph *G, uint64_t v)
{
    return G->rows + v * (G->R + 1) + 1;
}

/* beamSearch.h:43-171.  `self` is the id whose row is the query (build-time
 * search skips it, beamSearch.h:79 same_as); pass 0xFFFFFFFF for external
 * queries.  X / xq non-NULL: filtering mode (track_exact_dist): every pop
 * that lands in the visited list records its exact distance in a parallel
 * array moved with it (beamSearch.h:95-117); at the end the visited list
 * takes those distances and is re-sorted (163-168).  *exact_cmps counts
 * them. */
static uint64_t beam_search_x(cell_t *c, const nbpgo_graph *G, const uint8_t *q, uint32_t self,
                              const nbpgo_points *P, const nbpgo_qparams *qp,
                              const nbpgo_points *X, const uint8_t *xq, uint64_t *exact_cmps)
{
    const uint32_t L = qp->L;
    const uint64_t limit = qp->limit;
    uint64_t cmps = c->frontier_size;
    long remain = (long)c->frontier_size;
    long offset = 0;
    uint64_t cand_n = 0;

    while (remain > offset && c->num_visited < limit) {
        nbpgo_iddist cur = c->unvisited[offset];
        uint32_t deg = g_deg(G, cur.id);
        const uint32_t *nb = g_nbrs(G, cur.id);
        long num_elts = deg < qp->degree_limit ? (long)deg : (long)qp->degree_limit;
        uint64_t nf = 0;
        for (long i = 0; i < num_elts; ++i) {
            uint32_t a = nb[i];
            if (has_been_seen(c, a) || a == self)
                continue;
            c->filtered[nf++] = a;
        }
        for (uint64_t i = 0; i < nf; ++i) {
            uint32_t a = c->filtered[i];
            c->candidates[cand_n].id = a;
            c->candidates[cand_n].dist = pdist(P, a, q);
            ++cand_n;
            ++cmps;
        }
        /* beamSearch.h:95-117: sorted insertion into visited, cap max_visited */
        uint64_t lo = 0, hi = c->visited_size;
        while (lo < hi) { /* upper_bound */
            uint64_t mid = (lo + hi) / 2;
            if (less_id(&cur, &c->visited[mid]))
                hi = mid;
            else
                lo = mid + 1;
        }
        uint64_t pos = lo;
        c->num_visited++;
        if (c->visited_size < c->max_visited)
            c->visited_size++;
        if (pos < c->visited_size - 1) {
            memmove(c->visited + pos + 1, c->visited + pos,
                    sizeof(nbpgo_iddist) * (c->visited_size - 1 - pos));
            if (X)
                memmove(c->tmp_visited + pos + 1, c->tmp_visited + pos,
                        sizeof(float) * (c->visited_size - 1 - pos));
        }
        if (pos <= c->visited_size - 1) {
            c->visited[pos] = cur;
            if (X) {
                c->tmp_visited[pos] = pdist(X, cur.id, xq);
                ++*exact_cmps;
            }
        }

        /* beamSearch.h:119-125: delayed merge */
        if (cand_n == 0 ||
            (limit >= 2 * (uint64_t)L && cand_n < (uint64_t)(L / 8) && offset + 1 < remain)) {
            offset++;
            continue;
        }
        offset = 0;

        /* beamSearch.h:128-130: sort + unique by id */
        qsort(c->candidates, cand_n, sizeof(nbpgo_iddist), cmp_iddist);
        uint64_t cu = 1;
        for (uint64_t r = 1; r < cand_n; ++r)
            if (c->candidates[r].id != c->candidates[cu - 1].id)
                c->candidates[cu++] = c->candidates[r];

        /* beamSearch.h:132-135: set_union(frontier, candidates) */
        uint64_t i1 = 0, i2 = 0, o = 0;
        while (i1 < c->frontier_size) {
            if (i2 == cu) {
                while (i1 < c->frontier_size)
                    c->new_frontier[o++] = c->frontier[i1++];
                break;
            }
            if (less_id(&c->frontier[i1], &c->candidates[i2]))
                c->new_frontier[o++] = c->frontier[i1++];
            else if (less_id(&c->candidates[i2], &c->frontier[i1]))
                c->new_frontier[o++] = c->candidates[i2++];
            else {
                c->new_frontier[o++] = c->frontier[i1++];
                ++i2;
            }
        }
        while (i2 < cu)
            c->new_frontier[o++] = c->candidates[i2++];
        cand_n = 0;
        uint64_t nfs = o < L ? o : L;

        /* beamSearch.h:140-149: cut against the k-th entry */
        if (qp->k > 0 && nfs > qp->k) {
            nbpgo_iddist thr;
            thr.id = 0;
            thr.dist = (float)(qp->cut * (double)c->new_frontier[qp->k].dist);
            uint64_t a = 0, b = nfs;
            while (a < b) { /* upper_bound */
                uint64_t mid = (a + b) / 2;
                if (less_id(&thr, &c->new_frontier[mid]))
                    b = mid;
                else
                    a = mid + 1;
            }
            uint32_t ub = (uint32_t)a, fs = (uint32_t)c->frontier_size;
            nfs = ub > fs ? ub : fs;
        }
        memcpy(c->frontier, c->new_frontier, sizeof(nbpgo_iddist) * nfs);
        c->frontier_size = nfs;

        /* beamSearch.h:154-160: unvisited = frontier \ visited */
        uint64_t fl = c->frontier_size < L ? c->frontier_size : L;
        uint64_t f = 0, v = 0, r = 0;
        while (f < fl && v < c->visited_size) {
            if (less_id(&c->frontier[f], &c->visited[v]))
                c->unvisited[r++] = c->frontier[f++];
            else {
                if (!less_id(&c->visited[v], &c->frontier[f]))
                    ++f;
                ++v;
            }
        }
        while (f < fl)
            c->unvisited[r++] = c->frontier[f++];
        remain = (long)r;
    }
    if (X) {
        for (uint64_t i = 0; i < c->visited_size; ++i)
            c->visited[i].dist = c->tmp_visited[i];
        qsort(c->visited, c->visited_size, sizeof(nbpgo_iddist), cmp_iddist);
    }
    return cmps;
}

static uint64_t beam_search(cell_t *c, const nbpgo_graph *G, const uint8_t *q, uint32_t self,
                            const nbpgo_points *P, const nbpgo_qparams *qp)
{
    return beam_search_x(c, G, q, self, P, qp, NULL, NULL, NULL);
}

int nbpgo_search(const nbpgo_graph *G, const nbpgo_points *base, const nbpgo_points *queries,
                 const uint32_t *start_points, uint32_t sp_num, const nbpgo_qparams *qp,
                 uint32_t *out_ids, float *out_dists, uint32_t *out_stats, int nthreads)
{
    if (qp->k > qp->L) /* beamSearch.h:268-273 */
        return -1;
    if (sp_num > qp->L) /* query_params.h:206-210 */
        return -1;
    for (uint32_t s = 0; s < sp_num; ++s)
        if (start_points[s] >= base->n)
            return -1;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    int err = 0;
#pragma omp parallel
    {
        cell_t c;
        int ok = cell_alloc(&c, qp->L, G->R > qp->degree_limit ? G->R : qp->degree_limit,
                            qp->max_visited) == 0;
        if (!ok) {
#pragma omp atomic write
            err = -2;
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)queries->n; ++i) {
            if (!ok)
                continue;
            const uint8_t *q = row_of(queries, (uint64_t)i);
            cell_clear(&c);
            init_starting_points(&c, base, q, start_points, sp_num);
            uint64_t cmps = beam_search(&c, G, q, 0xFFFFFFFFu, base, qp);
            for (uint32_t j = 0; j < qp->k; ++j) {
                size_t o = (size_t)i * qp->k + j;
                if (j < c.frontier_size) {
                    out_ids[o] = c.frontier[j].id;
                    if (out_dists)
                        out_dists[o] = c.frontier[j].dist;
                } else {
                    out_ids[o] = 0xFFFFFFFFu;
                    if (out_dists)
                        out_dists[o] = INFINITY;
                }
            }
            if (out_stats) {
                out_stats[2 * i] = (uint32_t)c.num_visited;
                out_stats[2 * i + 1] = (uint32_t)cmps;
            }
        }
        if (ok)
            cell_free(&c);
    }
    return err;
}

/* qsearchAll with full-precision rows beside the quantized ones
 * (beam_search_rerank, beamSearch.h:183-247): rerank_factor > 0 re-scores
 * frontier[0:min(k * rerank_factor, fs)] with the raw rows and sorts by
 * (dist, id) (221-236); rerank_factor = 0 is the filtering mode
 * (beam_search_x with the raw rows), returning the first
 * min(k, fs, visited) of the exactly re-sorted visited list (217, 239-246). */
int nbpgo_search_exact(const nbpgo_graph *G, const nbpgo_points *base,
                       const nbpgo_points *queries, const nbpgo_points *raw_base,
                       const nbpgo_points *raw_queries, const uint32_t *start_points,
                       uint32_t sp_num, const nbpgo_qparams *qp, uint32_t rerank_factor,
                       uint32_t *out_ids, float *out_dists, uint32_t *out_stats, int nthreads)
{
    if (qp->k > qp->L || sp_num > qp->L || raw_base->n != base->n ||
        raw_queries->n != queries->n)
        return -1;
    for (uint32_t s = 0; s < sp_num; ++s)
        if (start_points[s] >= base->n)
            return -1;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    int err = 0;
#pragma omp parallel
    {
        cell_t c;
        nbpgo_iddist *pts = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * ((size_t)qp->L + 1));
        int ok = cell_alloc(&c, qp->L, G->R > qp->degree_limit ? G->R : qp->degree_limit,
                            qp->max_visited) == 0 && pts;
        if (!ok) {
#pragma omp atomic write
            err = -2;
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)queries->n; ++i) {
            if (!ok)
                continue;
            const uint8_t *q = row_of(queries, (uint64_t)i);
            const uint8_t *xq = row_of(raw_queries, (uint64_t)i);
            cell_clear(&c);
            init_starting_points(&c, base, q, start_points, sp_num);
            uint64_t exact = 0, cmps;
            if (rerank_factor == 0)
                cmps = beam_search_x(&c, G, q, 0xFFFFFFFFu, base, qp, raw_base, xq, &exact);
            else
                cmps = beam_search(&c, G, q, 0xFFFFFFFFu, base, qp);
            cmps += exact;
            if (c.visited_size > c.frontier_size)
                c.visited_size = c.frontier_size;
            const nbpgo_iddist *res;
            uint64_t kk;
            if (rerank_factor) {
                uint64_t nc = (uint64_t)qp->k * rerank_factor;
                if (nc > c.frontier_size)
                    nc = c.frontier_size;
                for (uint64_t j = 0; j < nc; ++j) {
                    pts[j].id = c.frontier[j].id;
                    pts[j].dist = pdist(raw_base, c.frontier[j].id, xq);
                }
                qsort(pts, nc, sizeof(nbpgo_iddist), cmp_iddist);
                res = pts;
                kk = nc < qp->k ? nc : qp->k;
            } else {
                res = c.visited;
                kk = c.visited_size < qp->k ? c.visited_size : qp->k;
            }
            for (uint32_t j = 0; j < qp->k; ++j) {
                size_t o = (size_t)i * qp->k + j;
                out_ids[o] = j < kk ? res[j].id : 0xFFFFFFFFu;
                if (out_dists)
                    out_dists[o] = j < kk ? res[j].dist : INFINITY;
            }
            if (out_stats) {
                out_stats[2 * i] = (uint32_t)c.num_visited;
                out_stats[2 * i + 1] = (uint32_t)cmps;
            }
        }
        if (ok)
            cell_free(&c);
        free(pts);
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* std::shuffle(first, last, std::mt19937(seed)), libstdc++ (GCC 11)     */
/* ------------------------------------------------------------------ */

typedef struct {
    uint32_t mt[624];
    int idx;
} mt19937_t;

static void mt_seed(mt19937_t *g, uint32_t s)
{
    g->mt[0] = s;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt_next(mt19937_t *g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* uniform_int_distribution<unsigned long>{0, hi}(g) for a 32-bit engine:
 * Lemire's nearly-divisionless reduction with a 64-bit product */
static uint64_t uniform_upto(mt19937_t *g, uint64_t hi)
{
    if (hi == 0xFFFFFFFFull)
        return mt_next(g);
    uint32_t range = (uint32_t)(hi + 1);
    uint64_t prod = (uint64_t)mt_next(g) * range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        uint32_t thr = (uint32_t)(-range) % range;
        while (low < thr) {
            prod = (uint64_t)mt_next(g) * range;
            low = (uint32_t)prod;
        }
    }
    return prod >> 32;
}

void nbpgo_shuffle_mt19937(uint32_t *a, uint64_t n, uint32_t seed)
{
    if (n == 0)
        return;
    mt19937_t g;
    mt_seed(&g, seed);
    const uint64_t rng = 0xFFFFFFFFull;
    if (rng / n >= n) {
        /* two swap positions per draw */
        uint64_t i = 1;
        if (n % 2 == 0) {
            uint64_t j = uniform_upto(&g, 1);
            uint32_t t = a[i];
            a[i] = a[j];
            a[j] = t;
            ++i;
        }
        while (i != n) {
            uint64_t b0 = i + 1, b1 = i + 2;
            uint64_t x = uniform_upto(&g, b0 * b1 - 1);
            uint64_t p0 = x / b1, p1 = x % b1;
            uint32_t t = a[i];
            a[i] = a[p0];
            a[p0] = t;
            ++i;
            t = a[i];
            a[i] = a[p1];
            a[p1] = t;
            ++i;
        }
        return;
    }
    for (uint64_t i = 1; i != n; ++i) {
        uint64_t j = uniform_upto(&g, i);
        uint32_t t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
}

uint64_t nbpgo_sp_stream(uint64_t seed, uint64_t counter)
{
    /* splitmix64 of (seed, counter) */
    uint64_t z = seed + (counter + 1) * UINT64_C(0x9e3779b97f4a7c15);
    z = (z ^ (z >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94d049bb133111eb);
    return z ^ (z >> 31);
}

/* ------------------------------------------------------------------ */
/* Build: nsw.h:110-663, merge_edges.h:22-148, dfs.h:62-253             */
/* ------------------------------------------------------------------ */

enum { POP_FIXED_MIN = 1, POP_FIXED_MAX, POP_CAL, POP_PRUNE };
enum { SP_SPLIT, SP_MERGE, SP_DEFAULT, SP_RAND };

typedef struct {
    uint32_t src;    /* inserted node */
    uint32_t nprune; /* pruned list size */
    uint32_t nvis;   /* visited list size */
} insert_result_t;

typedef struct {
    uint32_t target;
    nbpgo_iddist e; /* (source, dist) */
} edge_t;

static int cmp_edge(const void *x, const void *y)
{
    const edge_t *a = (const edge_t *)x, *b = (const edge_t *)y;
    if (a->target != b->target)
        return a->target < b->target ? -1 : 1;
    return cmp_iddist(&a->e, &b->e);
}

/* nsw.h:297-324 */
static void alpha_cos(int mode, uint32_t index, const float *popular, float minPop, float maxPop,
                      float minA, float maxA, float minC, float maxC, float *alpha, float *cosv)
{
    if (mode == POP_FIXED_MIN) {
        *alpha = maxA;
        *cosv = maxC;
    } else if (mode == POP_PRUNE) {
        float cof = (popular[index] - maxPop) / (minPop - maxPop);
        *alpha = minA + cof * (maxA - minA);
        *cosv = minC + cof * (maxC - minC);
    } else {
        *alpha = 1.0f;
        *cosv = minC;
    }
}

/* merge_union (utilities.h:128-165) over (dist,id) order */
static uint64_t merge_union(const nbpgo_iddist *a, uint64_t na, const nbpgo_iddist *b, uint64_t nb,
                            nbpgo_iddist *out, uint64_t cap)
{
    uint64_t i = 0, j = 0, k = 0;
    while (i < na && j < nb && k < cap) {
        if (less_id(&a[i], &b[j]))
            out[k++] = a[i++];
        else if (less_id(&b[j], &a[i]))
            out[k++] = b[j++];
        else {
            out[k++] = a[i++];
            ++j;
        }
    }
    while (i < na && k < cap)
        out[k++] = a[i++];
    while (j < nb && k < cap)
        out[k++] = b[j++];
    return k;
}

static void set_row(nbpgo_graph *G, uint32_t v, const nbpgo_iddist *l, uint64_t n)
{
    uint32_t *r = G->rows + (uint64_t)v * (G->R + 1);
    r[0] = (uint32_t)n;
    for (uint64_t i = 0; i < n; ++i)
        r[1 + i] = l[i].id;
}

typedef struct {
    const nbpgo_points *P;
    nbpgo_graph *G;
    const nbpgo_build_params *bp;
    float *popular;
    float minPop, maxPop;
    uint64_t sp_counter;
} builder_t;

static void batch_insert(builder_t *B, const uint32_t *inserts, uint64_t m, int pop_mode,
                         int sp_mode, int nthreads)
{
    const nbpgo_points *P = B->P;
    nbpgo_graph *G = B->G;
    const nbpgo_build_params *bp = B->bp;
    const uint32_t R = G->R, L = bp->L, VL = bp->VL;
    const uint32_t top_degree = bp->top_degree == 0 ? R : bp->top_degree;
    const uint64_t N = G->n;
    /* nsw.h:288-295 */
    uint64_t maxb = (uint64_t)(0.02 * (double)(float)N);
    if (maxb > 1000000)
        maxb = 1000000;
    if (maxb == 0)
        maxb = 1;
    const uint32_t nsp = (sp_mode == SP_MERGE) ? bp->num_split : 1;

    nbpgo_iddist *prune_res = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * maxb * (VL + R + 1));
    nbpgo_iddist *vis_res = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * maxb * (VL + 1));
    insert_result_t *res = (insert_result_t *)malloc(sizeof(insert_result_t) * maxb);
    uint32_t *sp = (uint32_t *)malloc(sizeof(uint32_t) * maxb * nsp);
    edge_t *edges = (edge_t *)malloc(sizeof(edge_t) * maxb * R + 1);
    uint64_t *gstart = (uint64_t *)malloc(sizeof(uint64_t) * (maxb * R + 2));

    if (pop_mode == POP_CAL)
        memset(B->popular, 0, sizeof(float) * N);

    uint64_t inc = 0, count = 0;
    while (count < m) {
        /* nsw.h:30-47 next_batch_range, base 2 */
        uint64_t floor_, ceil_;
        double p2 = pow(2.0, (double)inc);
        if (p2 <= (double)maxb) {
            floor_ = (uint64_t)p2 - 1;
            ceil_ = (uint64_t)pow(2.0, (double)(inc + 1)) - 1;
            if (ceil_ > m)
                ceil_ = m;
            count = ceil_;
        } else {
            floor_ = count;
            ceil_ = count + maxb < m ? count + maxb : m;
            count = count + (ceil_ - floor_);
        }
        const uint64_t bs = ceil_ - floor_;
        /* start points (nsw.h:567-616), filled serially */
        for (uint64_t j = 0; j < bs; ++j) {
            uint32_t index = inserts[floor_ + j];
            if (sp_mode == SP_SPLIT)
                sp[j * nsp] = index % bp->num_split;
            else if (sp_mode == SP_DEFAULT)
                sp[j * nsp] = 0;
            else if (sp_mode == SP_RAND)
                sp[j * nsp] = (uint32_t)(m - index - 1);
            else {
                for (uint32_t s = 0; s < nsp; ++s) {
                    uint32_t rid = (uint32_t)(nbpgo_sp_stream(bp->sp_seed, B->sp_counter++) % N);
                    rid = rid / bp->num_split * bp->num_split;
                    uint32_t v = rid + s;
                    if (v >= N)
                        v -= bp->num_split;
                    sp[j * nsp + s] = v;
                }
            }
        }
        /* batch_search (nsw.h:380-414) + searchNghAndPrune (nsw.h:122-175) */
        if (nthreads > 0)
            omp_set_num_threads(nthreads);
#pragma omp parallel
        {
            cell_t c;
            cell_alloc(&c, L, R, VL);
#pragma omp for schedule(dynamic, 1)
            for (int64_t j = 0; j < (int64_t)bs; ++j) {
                uint32_t index = inserts[floor_ + j];
                float alpha, cosv;
                alpha_cos(pop_mode, index, B->popular, B->minPop, B->maxPop, (float)bp->min_alpha,
                          (float)bp->max_alpha, (float)bp->min_cos_angle, (float)bp->max_cos_angle,
                          &alpha, &cosv);
                nbpgo_qparams qp = {0, 0.0, L, bp->limit, R, VL};
                const uint8_t *q = row_of(P, index);
                cell_clear(&c);
                init_starting_points(&c, P, q, sp + j * nsp, nsp);
                beam_search(&c, G, q, index, P, &qp);
                /* existing neighbours unseen by the filter (nsw.h:150-161) */
                uint32_t deg = g_deg(G, index);
                const uint32_t *nb = g_nbrs(G, index);
                c.frontier_size = 0;
                for (uint32_t i = 0; i < deg; ++i)
                    if (!has_been_seen(&c, nb[i])) {
                        c.frontier[c.frontier_size].id = nb[i];
                        c.frontier[c.frontier_size].dist = pdist(P, nb[i], q);
                        c.frontier_size++;
                    }
                uint64_t n = merge_union(c.frontier, c.frontier_size, c.visited, c.visited_size,
                                         c.copy_visited, VL);
                n = nbpgo_sort_candidates(index, c.copy_visited, n);
                nbpgo_robust_prune(index, c.copy_visited, &n, P, R, (double)alpha, (double)cosv, 1);
                memcpy(prune_res + (uint64_t)j * (VL + R + 1), c.copy_visited,
                       sizeof(nbpgo_iddist) * n);
                memcpy(vis_res + (uint64_t)j * (VL + 1), c.visited,
                       sizeof(nbpgo_iddist) * c.visited_size);
                res[j].src = index;
                res[j].nprune = (uint32_t)n;
                res[j].nvis = (uint32_t)c.visited_size;
            }
            cell_free(&c);
        }
        if (pop_mode == POP_CAL) /* nsw.h:220-230 */
            for (uint64_t j = 0; j < bs; ++j)
                for (uint32_t i = 0; i < res[j].nvis; ++i)
                    B->popular[vis_res[j * (VL + 1) + i].id] += 1.0f;
        /* forward edges (nsw.h:416-429) */
        for (uint64_t j = 0; j < bs; ++j)
            set_row(G, res[j].src, prune_res + j * (VL + R + 1), res[j].nprune);
        /* reverse edges (nsw.h:431-444, merge_edges.h:53-81) */
        uint64_t ne = 0;
        for (uint64_t j = 0; j < bs; ++j)
            for (uint32_t i = 0; i < res[j].nprune; ++i) {
                edges[ne].target = prune_res[j * (VL + R + 1) + i].id;
                edges[ne].e.id = res[j].src;
                edges[ne].e.dist = prune_res[j * (VL + R + 1) + i].dist;
                ++ne;
            }
        qsort(edges, ne, sizeof(edge_t), cmp_edge);
        uint64_t ng = 0;
        for (uint64_t i = 0; i < ne; ++i)
            if (i == 0 || edges[i].target != edges[i - 1].target)
                gstart[ng++] = i;
        gstart[ng] = ne;
        /* write-back (nsw.h:463-485, 489-511) */
#pragma omp parallel
        {
            nbpgo_iddist *cand = (nbpgo_iddist *)malloc(sizeof(nbpgo_iddist) * (VL + R + 1));
#pragma omp for schedule(dynamic, 16)
            for (int64_t g = 0; g < (int64_t)ng; ++g) {
                uint64_t b = gstart[g], e = gstart[g + 1];
                if (e - b > VL)
                    e = b + VL;
                uint32_t target = edges[b].target;
                uint64_t cs = 0;
                for (uint64_t i = b; i < e; ++i)
                    cand[cs++] = edges[i].e;
                uint64_t pre = cs;
                /* add_neighbors_without_repeats (nsw.h:358-378) */
                uint32_t deg = g_deg(G, target);
                const uint32_t *nb = g_nbrs(G, target);
                for (uint32_t i = 0; i < deg; ++i) {
                    int exists = 0;
                    for (uint64_t t = 0; t < cs; ++t)
                        if (cand[t].id == nb[i]) {
                            exists = 1;
                            break;
                        }
                    if (exists || cs >= (uint64_t)VL + R)
                        continue;
                    cand[cs++].id = nb[i];
                }
                /* eval_dist_of_new_neighbors (nsw.h:329-355) */
                const uint8_t *tq = row_of(P, target);
                uint64_t end = pre + top_degree < cs ? pre + top_degree : cs;
                for (uint64_t t = pre; t < end; ++t)
                    cand[t].dist = pdist(P, cand[t].id, tq);
                if (end < cs) {
                    float a = cand[end - 1].dist;
                    float bb = pdist(P, cand[cs - 1].id, tq);
                    cand[cs - 1].dist = bb;
                    float delta = (a - bb) / (float)(cs - end);
                    for (uint64_t t = end; t < cs - 1; ++t)
                        cand[t].dist = a + (float)(t - end + 1) * delta;
                }
                if (cs <= R)
                    qsort(cand, cs, sizeof(nbpgo_iddist), cmp_iddist);
                else {
                    cs = dedupe_by_index(cand, cs);
                    qsort(cand, cs, sizeof(nbpgo_iddist), cmp_iddist);
                    if (cs > R)
                        cs = R;
                }
                set_row(G, target, cand, cs);
            }
            free(cand);
        }
        inc += 1;
    }

    if (pop_mode == POP_CAL) { /* nsw.h:255-278 */
        float *tmp = (float *)calloc(N, sizeof(float));
        for (uint64_t i = 0; i < N; ++i) {
            tmp[i] = B->popular[i];
            uint32_t deg = g_deg(G, i);
            const uint32_t *nb = g_nbrs(G, i);
            for (uint32_t j = 0; j < deg; ++j)
                tmp[nb[j]] += tmp[i];
        }
        memcpy(B->popular, tmp, sizeof(float) * N);
        free(tmp);
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (uint64_t i = 0; i < N; ++i) {
            B->popular[i] = (float)log((double)(B->popular[i] + 1.0f));
            if (B->popular[i] < mn)
                mn = B->popular[i];
            if (B->popular[i] > mx)
                mx = B->popular[i];
        }
        B->minPop = mn;
        B->maxPop = mx;
    }
    free(prune_res);
    free(vis_res);
    free(res);
    free(sp);
    free(edges);
    free(gstart);
}

uint64_t nbpgo_repair_isolated(nbpgo_graph *G, uint32_t spn)
{
    /* dfs.h:101-137 BFS from the entry, then dfs.h:192-215 */
    uint64_t N = G->n;
    int *flag = (int *)calloc(N, sizeof(int));
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * (N + 1));
    uint64_t qh = 0, qt = 0;
    flag[spn] = 1;
    queue[qt++] = spn;
    while (qh < qt) {
        uint32_t v = queue[qh++];
        uint32_t deg = g_deg(G, v);
        const uint32_t *nb = g_nbrs(G, v);
        for (uint32_t j = 0; j < deg; ++j)
            if (!flag[nb[j]]) {
                flag[nb[j]] = 1;
                queue[qt++] = nb[j];
            }
    }
    uint64_t cnt = 0;
    for (uint64_t i = 0; i < N; ++i) {
        if (flag[i])
            continue;
        ++cnt;
        /* try_repair_isolated_node (dfs.h:62-98) */
        uint32_t deg = g_deg(G, i);
        uint32_t nb[4096];
        memcpy(nb, g_nbrs(G, i), sizeof(uint32_t) * deg);
        int done = 0;
#define ADD_EDGE(from)                                                                             \
    do {                                                                                           \
        uint32_t *r_ = G->rows + (uint64_t)(from) * (G->R + 1);                                    \
        if (r_[0] < G->R) {                                                                        \
            int ex_ = 0;                                                                           \
            for (uint32_t k_ = 0; k_ < r_[0]; ++k_)                                                \
                if (r_[1 + k_] == (uint32_t)i)                                                     \
                    ex_ = 1;                                                                       \
            if (!ex_) {                                                                            \
                r_[1 + r_[0]] = (uint32_t)i;                                                       \
                r_[0]++;                                                                           \
            }                                                                                      \
            done = 1;                                                                              \
        }                                                                                          \
    } while (0)
        for (uint32_t j = 0; j < deg && !done; ++j)
            if (flag[nb[j]] == 1)
                ADD_EDGE(nb[j]);
        for (uint32_t j = 0; j < deg && !done; ++j)
            ADD_EDGE(nb[j]);
        if (!done && flag[spn] == 1)
            ADD_EDGE(spn);
        for (uint64_t r = 0; r < N && !done; ++r)
            if (flag[r] == 1)
                ADD_EDGE(r);
#undef ADD_EDGE
    }
    free(flag);
    free(queue);
    return cnt;
}

int nbpgo_build(const nbpgo_points *P, nbpgo_graph *G, const nbpgo_build_params *bp,
                uint64_t *out_isolated, int nthreads)
{
    if (G->n != P->n || G->R > 4096)
        return -1;
    const uint64_t N = P->n;
    builder_t B;
    memset(&B, 0, sizeof(B));
    B.P = P;
    B.G = G;
    B.bp = bp;
    B.popular = (float *)calloc(N, sizeof(float));
    B.maxPop = 1.0f;
    B.minPop = 0.0f;
    uint32_t *ins = (uint32_t *)malloc(sizeof(uint32_t) * N);
    for (uint64_t i = 0; i < N; ++i)
        ins[i] = (uint32_t)i;
    int pop = POP_FIXED_MAX;
    if (bp->num_split >= 1) /* nsw.h:564-582 */
        for (int i = 0; i < 2; ++i) {
            nbpgo_shuffle_mt19937(ins, N, (uint32_t)i);
            batch_insert(&B, ins, N, pop, SP_SPLIT, nthreads);
        }
    for (uint32_t i = 0; i < bp->num_passes; ++i) { /* nsw.h:618-654 */
        nbpgo_shuffle_mt19937(ins, N, i);
        if (bp->num_passes >= 2 && i == bp->num_passes - 2 && bp->dynamic_prune)
            pop = POP_CAL;
        else if (i == bp->num_passes - 1)
            pop = (bp->dynamic_prune && bp->num_passes > 1) ? POP_PRUNE : POP_FIXED_MIN;
        int spm;
        if (i != bp->num_passes - 1)
            spm = bp->num_split > 1 ? SP_MERGE : SP_DEFAULT;
        else
            spm = bp->rand_sp ? SP_RAND : SP_DEFAULT;
        batch_insert(&B, ins, N, pop, spm, nthreads);
    }
    uint64_t iso = 0;
    if (bp->repair_isolate)
        iso = nbpgo_repair_isolated(G, 0);
    if (out_isolated)
        *out_isolated = iso;
    free(ins);
    free(B.popular);
    return 0;
}

/* ------------------------------------------------------------------ */
/* PCA compress: matrix.h:86-127                                        */
/* ------------------------------------------------------------------ */

void nbpgo_matmul(const float *x, uint64_t n, uint32_t d_in, const float *W, const float *bias,
                  uint32_t d_out, float *y, int nthreads)
{
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float *xr = x + (size_t)j * d_in;
        for (uint32_t o = 0; o < d_out; ++o) {
            const float *w = W + (size_t)o * d_in;
            float s = 0.0f;
            for (uint32_t i = 0; i < d_in; ++i)
                s += w[i] * xr[i];
            y[(size_t)j * d_out + o] = bias[o] - (-s);
        }
    }
}

void nbpgo_compress(const float *y, uint64_t n, uint32_t d, int is_signed, uint32_t chunk,
                    uint8_t *out, int nthreads)
{
    const int maxn = 255;
    const int shift = 8 - (int)(8 / chunk);
    const int delta = is_signed ? 128 : 0;
    const int mask = (1 << (8 / chunk)) - 1;
    const uint32_t od = d / chunk;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)n; ++p) {
        const float *v = y + (size_t)p * d;
        for (uint32_t i = 0; i < od; ++i) {
            uint8_t r = 0;
            for (uint32_t j = 0; j < chunk; ++j) {
                float f = v[(i + 1) * chunk - 1 - j];
                r = (uint8_t)(r << (8 / chunk));
                int ti = (int)((double)(f + 1.0f) * 0.5 * (double)maxn);
                ti = ti < 0 ? 0 : (ti > maxn ? maxn : ti);
                ti = (ti - delta) >> shift;
                r = (uint8_t)(r + (uint8_t)(ti & mask));
            }
            out[(size_t)p * od + i] = r;
        }
    }
}

/* ------------------------------------------------------------------ */
/* recall with ties: check_nn_recall.h:126-167                           */
/* ------------------------------------------------------------------ */

double nbpgo_recall(const uint32_t *result_ids, uint32_t res_stride, uint64_t nq,
                    const uint32_t *gt_ids, uint32_t gt_k, const nbpgo_points *base,
                    const nbpgo_points *queries, uint32_t k)
{
    uint64_t correct = 0;
    for (uint64_t i = 0; i < nq; ++i) {
        const uint32_t *g = gt_ids + i * gt_k;
        const uint8_t *q = row_of(queries, i);
        float last = pdist(base, g[k - 1], q);
        uint32_t nt = k;
        for (uint32_t l = k; l < gt_k; ++l) {
            if (pdist(base, g[l], q) == last)
                ++nt;
            else
                break;
        }
        const uint32_t *r = result_ids + i * res_stride;
        for (uint32_t l = 0; l < nt; ++l) {
            for (uint32_t t = 0; t < k; ++t)
                if (r[t] == g[l]) {
                    ++correct;
                    break;
                }
        }
    }
    return (double)correct / ((double)k * (double)nq);
}
