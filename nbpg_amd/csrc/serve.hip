// serve.hip -- continuous-batching request loop over a resident index.
//
// Requests (1..N query rows, each request with its own L and k) are queued
// by nbpg_server_submit from any thread.  A scheduler thread owns `slots`
// in-flight batches, each with its own HIP stream, work counter and
// device / pinned-host buffers sized at creation.  Whenever a slot is free
// it admits queued requests in arrival order until the slot's query capacity
// (the memory admitted per batch) is full, stages their rows through pinned
// memory, and launches one persistent FAST search over the batch
// (per-query L and k, the uneven-batch kernel) followed by the result copy
// and an event.  Batches on different slots overlap on the device; a request
// that arrives while others run is admitted as soon as a slot frees.
//
// Every query is searched by its own wavefront with its own LDS filter and
// beam, and nothing in that search depends on the other queries of its
// batch, so a request's results are the same whether it ran alone or in a
// mix (only L, k and the launch shape matter).
//
// Latency = completion seen by the scheduler - submit time, per request.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "search_args.h"

namespace nbpg {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    uint32_t nq, L, k;
    std::vector<uint8_t> rows; // nq x row_stride
    std::vector<uint32_t> ids; // nq x k
    std::vector<float> dists;
    Clock::time_point t_submit, t_done;
    bool done = false;
    bool waited = false; // a thread is inside nbpg_server_wait for it
    int status = NBPG_OK;
    std::string error;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint8_t *dq = nullptr;
    uint32_t *dL = nullptr, *dk = nullptr, *dids = nullptr, *counter = nullptr;
    float *dd = nullptr;
    uint8_t *hq = nullptr;
    uint32_t *hL = nullptr, *hk = nullptr, *hids = nullptr;
    float *hd = nullptr;
    std::vector<Request *> reqs;
    uint32_t nq = 0, out_k = 0;
    bool busy = false;
};

} // namespace nbpg

struct nbpg_server {
    const nbpg_index *ix = nullptr;
    nbpg_server_config cfg;
    uint32_t cap = 0;    // queries per batch
    uint32_t stride = 0; // bytes per staged query row
    int device = 0;
    std::vector<nbpg::Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<nbpg::Request *> pending;
    std::unordered_map<uint64_t, nbpg::Request *> live;
    uint64_t next_ticket = 1;
    uint32_t waiters = 0; // threads inside nbpg_server_wait
    bool stop = false;
    std::thread worker;
    // statistics since creation or the last reset
    std::vector<float> lat_ms;
    uint64_t batches = 0, batch_queries = 0, done_queries = 0;
};

namespace nbpg {

static void free_slot(Slot &s)
{
    hipFree(s.dq);
    hipFree(s.dL);
    hipFree(s.dk);
    hipFree(s.dids);
    hipFree(s.dd);
    hipFree(s.counter);
    hipHostFree(s.hq);
    hipHostFree(s.hL);
    hipHostFree(s.hk);
    hipHostFree(s.hids);
    hipHostFree(s.hd);
    if (s.ev)
        hipEventDestroy(s.ev);
    if (s.stream)
        hipStreamDestroy(s.stream);
    s = Slot();
}

static int alloc_slot(Slot &s, uint32_t cap, uint32_t stride, uint32_t max_k)
{
    const size_t q = (size_t)cap * stride, r = (size_t)cap * max_k;
    HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    HIP_TRY(hipMalloc((void **)&s.dq, q));
    HIP_TRY(hipMalloc((void **)&s.dL, cap * 4ull));
    HIP_TRY(hipMalloc((void **)&s.dk, cap * 4ull));
    HIP_TRY(hipMalloc((void **)&s.dids, r * 4));
    HIP_TRY(hipMalloc((void **)&s.dd, r * 4));
    HIP_TRY(hipMalloc((void **)&s.counter, 256));
    HIP_TRY(hipMemset(s.counter, 0, 256)); // work counters start at zero (next_query)
    HIP_TRY(hipHostMalloc((void **)&s.hq, q, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.hL, cap * 4ull, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.hk, cap * 4ull, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.hids, r * 4, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&s.hd, r * 4, hipHostMallocDefault));
    return NBPG_OK;
}

// caller holds the lock; stages and launches the slot's batch
static int launch_slot(nbpg_server *sv, Slot &s)
{
    uint32_t off = 0, Lmax = 0, kmax = 0;
    for (Request *r : s.reqs) {
        std::memcpy(s.hq + (size_t)off * sv->stride, r->rows.data(), (size_t)r->nq * sv->stride);
        for (uint32_t i = 0; i < r->nq; ++i) {
            s.hL[off + i] = r->L;
            s.hk[off + i] = r->k;
        }
        off += r->nq;
        Lmax = std::max(Lmax, r->L);
        kmax = std::max(kmax, r->k);
    }
    s.nq = off;
    s.out_k = kmax;
    HIP_TRY(hipMemcpyAsync(s.dq, s.hq, (size_t)off * sv->stride, hipMemcpyHostToDevice, s.stream));
    HIP_TRY(hipMemcpyAsync(s.dL, s.hL, off * 4ull, hipMemcpyHostToDevice, s.stream));
    HIP_TRY(hipMemcpyAsync(s.dk, s.hk, off * 4ull, hipMemcpyHostToDevice, s.stream));
    if (sv->cfg.seed_nearest)
        NBPG_TRY(start_select(sv->ix, s.dq, off, sv->stride, s.dids, kmax, s.stream));
    NBPG_TRY(search_fast(sv->ix, s.dq, off, sv->stride, sv->cfg.start_points[0],
                         sv->cfg.start_points, sv->cfg.sp_num, 0, 0, s.dL, s.dk, Lmax, kmax,
                         sv->ix->n, &sv->cfg.fast, s.dids, s.dd, nullptr, s.counter, s.stream,
                         2, nullptr, 0, false, sv->cfg.seed_nearest != 0));
    HIP_TRY(hipMemcpyAsync(s.hids, s.dids, (size_t)off * kmax * 4, hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipMemcpyAsync(s.hd, s.dd, (size_t)off * kmax * 4, hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipEventRecord(s.ev, s.stream));
    s.busy = true;
    return NBPG_OK;
}

static void finish_slot(nbpg_server *sv, Slot &s, int status, const char *err)
{
    const Clock::time_point now = Clock::now();
    uint32_t off = 0;
    for (Request *r : s.reqs) {
        r->status = status;
        if (status == NBPG_OK) {
            for (uint32_t i = 0; i < r->nq; ++i) {
                std::memcpy(r->ids.data() + (size_t)i * r->k, s.hids + (size_t)(off + i) * s.out_k,
                            r->k * 4ull);
                std::memcpy(r->dists.data() + (size_t)i * r->k, s.hd + (size_t)(off + i) * s.out_k,
                            r->k * 4ull);
            }
        } else {
            r->error = err ? err : "";
        }
        off += r->nq;
        r->t_done = now;
        r->done = true;
        sv->lat_ms.push_back(
            (float)std::chrono::duration<double, std::milli>(now - r->t_submit).count());
        sv->done_queries += r->nq;
    }
    s.reqs.clear();
    s.busy = false;
}

static void scheduler(nbpg_server *sv)
{
    hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        bool any_busy = false, progressed = false;
        for (Slot &s : sv->slots) {
            if (!s.busy)
                continue;
            const hipError_t e = hipEventQuery(s.ev);
            if (e == hipErrorNotReady) {
                any_busy = true;
                continue;
            }
            finish_slot(sv, s, e == hipSuccess ? NBPG_OK : NBPG_EHIP,
                        e == hipSuccess ? nullptr : hipGetErrorString(e));
            progressed = true;
        }
        if (progressed)
            sv->cv_done.notify_all();
        for (Slot &s : sv->slots) {
            if (s.busy || sv->pending.empty())
                continue;
            uint32_t nq = 0;
            while (!sv->pending.empty() && nq + sv->pending.front()->nq <= sv->cap) {
                s.reqs.push_back(sv->pending.front());
                nq += sv->pending.front()->nq;
                sv->pending.pop_front();
            }
            const int rc = launch_slot(sv, s);
            if (rc != NBPG_OK) {
                finish_slot(sv, s, rc, nbpg_last_error());
                sv->cv_done.notify_all();
                continue;
            }
            sv->batches++;
            sv->batch_queries += nq;
            any_busy = true;
        }
        if (sv->stop && !any_busy && sv->pending.empty())
            break;
        if (any_busy) {
            // poll the in-flight batches; new requests are picked up at once
            sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
        } else {
            sv->cv_work.wait(lk, [sv] { return sv->stop || !sv->pending.empty(); });
        }
    }
}

} // namespace nbpg

using namespace nbpg;

extern "C" {

int nbpg_server_create(nbpg_server **out, const nbpg_index *ix, const nbpg_server_config *cfg)
{
    NBPG_TRY(ensure_init());
    if (!out || !ix || !cfg)
        return set_error(NBPG_EINVAL, "null argument");
    if (cfg->slots < 1 || cfg->slots > 8 || cfg->max_batch_queries == 0 || cfg->max_L == 0 ||
        cfg->max_k == 0 || cfg->max_k > cfg->max_L || cfg->sp_num == 0 || cfg->sp_num > 64)
        return set_error(NBPG_EINVAL, "server config out of range");
    for (uint32_t i = 0; i < cfg->sp_num; ++i) {
        if (cfg->start_points[i] >= ix->n)
            return set_error(NBPG_EINVAL, "start point %u >= n", cfg->start_points[i]);
        for (uint32_t j = 0; j < i; ++j)
            if (cfg->start_points[j] == cfg->start_points[i])
                return set_error(NBPG_EINVAL, "start point %u repeated", cfg->start_points[i]);
    }
    if (cfg->sp_num > 1 && cfg->max_L > 4 * WAVE)
        return set_error(NBPG_EUNSUP, "several start points need max_L <= 256");
    if (cfg->seed_nearest && (cfg->max_L > 4 * WAVE || ix->n_seeds == 0 ||
                              !start_select_supported(ix)))
        return set_error(NBPG_EINVAL, "nearest-seed starts need max_L <= 256 and the index's "
                                      "seed rows (u8 / i8 rows of a multiple of 16 bytes)");
    nbpg_server *sv = new nbpg_server();
    sv->ix = ix;
    sv->cfg = *cfg;
    sv->device = g_dev.device;
    sv->stride = ix->stride; // staged rows padded like the index's own
    const uint64_t per_query = (uint64_t)sv->stride + 8 + 8ull * cfg->max_k;
    uint64_t cap = cfg->max_batch_queries;
    if (cfg->mem_budget)
        cap = std::min<uint64_t>(cap, cfg->mem_budget / cfg->slots / per_query);
    if (cap == 0) {
        delete sv;
        return set_error(NBPG_EINVAL, "memory budget admits no query per batch");
    }
    sv->cap = (uint32_t)std::min<uint64_t>(cap, 1u << 24);
    sv->slots.resize(cfg->slots);
    for (Slot &s : sv->slots) {
        const int rc = alloc_slot(s, sv->cap, sv->stride, cfg->max_k);
        if (rc != NBPG_OK) {
            for (Slot &t : sv->slots)
                free_slot(t);
            delete sv;
            return rc;
        }
    }
    sv->worker = std::thread(scheduler, sv);
    *out = sv;
    return NBPG_OK;
}

int nbpg_server_submit(nbpg_server *sv, const void *host_queries, uint32_t nq, uint64_t query_stride,
                       uint32_t L, uint32_t k, uint64_t *ticket)
{
    if (!ticket)
        return set_error(NBPG_EINVAL, "null argument");
    return nbpg_server_submit_many(sv, host_queries, 1, &nq, query_stride, L, k, ticket);
}

int nbpg_server_submit_many(nbpg_server *sv, const void *host_queries, uint32_t n,
                            const uint32_t *req_nq, uint64_t query_stride, uint32_t L, uint32_t k,
                            uint64_t *tickets)
{
    if (!sv || !host_queries || !tickets || !req_nq || n == 0)
        return set_error(NBPG_EINVAL, "null argument");
    if (L == 0 || L > sv->cfg.max_L || k == 0 || k > L || k > sv->cfg.max_k)
        return set_error(NBPG_EINVAL, "L = %u, k = %u (1 <= k <= L <= %u, k <= %u)", L, k,
                         sv->cfg.max_L, sv->cfg.max_k);
    if (query_stride < sv->ix->dims)
        return set_error(NBPG_EINVAL, "query stride too small");
    for (uint32_t i = 0; i < n; ++i) {
        if (req_nq[i] == 0)
            return set_error(NBPG_EINVAL, "request %u has no queries", i);
        if (req_nq[i] > sv->cap)
            return set_error(NBPG_EINVAL, "request of %u queries exceeds the batch capacity %u",
                             req_nq[i], sv->cap);
    }
    // rows copied outside the lock; all requests queued under one lock
    std::vector<Request *> rs(n);
    const uint8_t *src = (const uint8_t *)host_queries;
    for (uint32_t i = 0; i < n; ++i) {
        Request *r = new Request();
        r->nq = req_nq[i];
        r->L = L;
        r->k = k;
        r->rows.assign((size_t)r->nq * sv->stride, 0);
        for (uint32_t j = 0; j < r->nq; ++j)
            std::memcpy(r->rows.data() + (size_t)j * sv->stride, src + j * query_stride,
                        sv->ix->dims);
        src += (size_t)r->nq * query_stride;
        r->ids.resize((size_t)r->nq * k);
        r->dists.resize((size_t)r->nq * k);
        rs[i] = r;
    }
    {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->stop) {
            for (Request *r : rs)
                delete r;
            return set_error(NBPG_EINVAL, "server is shutting down");
        }
        const Clock::time_point now = Clock::now();
        for (uint32_t i = 0; i < n; ++i) {
            Request *r = rs[i];
            r->ticket = sv->next_ticket++;
            r->t_submit = now;
            sv->live[r->ticket] = r;
            sv->pending.push_back(r);
            tickets[i] = r->ticket;
        }
    }
    sv->cv_work.notify_one();
    return NBPG_OK;
}

} // extern "C"

namespace nbpg {

// nbpg_server_wait for one ticket; *entries = the request's nq * k (0 when
// the ticket is unknown)
static int wait_one(nbpg_server *sv, uint64_t ticket, uint32_t *host_ids, float *host_dists,
                    double *latency_ms, size_t *entries)
{
    *entries = 0;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->live.find(ticket);
    if (it == sv->live.end())
        return set_error(NBPG_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second;
    if (r->waited)
        return set_error(NBPG_EINVAL, "ticket %llu is already being waited on",
                         (unsigned long long)ticket);
    r->waited = true;
    sv->waiters++;
    sv->cv_done.wait(lk, [r] { return r->done; });
    // by key: submits while this thread slept may have rehashed the map
    sv->live.erase(ticket);
    sv->waiters--;
    if (sv->stop)
        sv->cv_done.notify_all(); // destroy waits for the last waiter to leave
    lk.unlock();
    int rc = r->status;
    *entries = r->ids.size();
    if (rc == NBPG_OK) {
        if (host_ids)
            std::memcpy(host_ids, r->ids.data(), r->ids.size() * 4);
        if (host_dists)
            std::memcpy(host_dists, r->dists.data(), r->dists.size() * 4);
        if (latency_ms)
            *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    } else {
        set_error(rc, "batch failed: %s", r->error.c_str());
    }
    delete r;
    return rc;
}

} // namespace nbpg

extern "C" {

int nbpg_server_wait(nbpg_server *sv, uint64_t ticket, uint32_t *host_ids, float *host_dists,
                     double *latency_ms)
{
    if (!sv)
        return set_error(NBPG_EINVAL, "null argument");
    size_t e = 0;
    return wait_one(sv, ticket, host_ids, host_dists, latency_ms, &e);
}

int nbpg_server_wait_many(nbpg_server *sv, uint32_t n, const uint64_t *tickets, uint32_t *host_ids,
                          float *host_dists, double *latency_ms)
{
    if (!sv || (n && !tickets))
        return set_error(NBPG_EINVAL, "null argument");
    int first = NBPG_OK;
    std::string msg;
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        size_t e = 0;
        const int rc = wait_one(sv, tickets[i], host_ids ? host_ids + off : nullptr,
                                host_dists ? host_dists + off : nullptr,
                                latency_ms ? latency_ms + i : nullptr, &e);
        if (rc != NBPG_OK && e == 0)
            return rc; // unknown ticket: the later requests' places are unknown
        if (rc != NBPG_OK && first == NBPG_OK) {
            first = rc;
            msg = nbpg_last_error();
        }
        off += e;
    }
    if (first != NBPG_OK)
        return set_error(first, "%s", msg.c_str());
    return NBPG_OK;
}

int nbpg_server_stats(nbpg_server *sv, int reset, nbpg_server_stats_t *st)
{
    if (!sv || !st)
        return set_error(NBPG_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(sv->mu);
    std::vector<float> v = sv->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) -> double {
        if (v.empty())
            return 0.0;
        const size_t i = std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5));
        return v[i];
    };
    st->requests = v.size();
    st->queries = sv->done_queries;
    st->batches = sv->batches;
    st->mean_batch_queries = sv->batches ? (double)sv->batch_queries / (double)sv->batches : 0.0;
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.empty() ? 0.0 : v.back();
    st->batch_capacity = sv->cap;
    if (reset) {
        sv->lat_ms.clear();
        sv->batches = sv->batch_queries = sv->done_queries = 0;
    }
    return NBPG_OK;
}

int nbpg_server_destroy(nbpg_server *sv)
{
    if (!sv)
        return NBPG_OK;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable())
        sv->worker.join();
    {
        // every request is done now; threads already blocked in
        // nbpg_server_wait take their results and leave before the
        // requests nobody waits for are freed
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.notify_all();
        sv->cv_done.wait(lk, [sv] { return sv->waiters == 0; });
        for (auto &kv : sv->live)
            delete kv.second;
        sv->live.clear();
    }
    for (Slot &s : sv->slots)
        free_slot(s);
    delete sv;
    return NBPG_OK;
}

} // extern "C"
