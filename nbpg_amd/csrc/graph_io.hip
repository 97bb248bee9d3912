// graph_io.hip -- the resident graph's file I/O and isolated-node repair,
// done on the device so a built index never round-trips its padded
// adjacency through host memory.
//
// * nbpg_index_save_graph / nbpg_index_load_graph: Graph::save / Graph(file)
//   (graph.h:169-221, 339-379): [u32 n][u32 R][u32 deg x n][ids, deg[i] per
//   node, node order].  The variable-length id stream is packed / unpacked
//   by kernels in node ranges of at most 2^26 ids; the host only moves the
//   packed bytes between the file and a staging buffer.
// * nbpg_index_repair: DFS::repair_isolate_node(G, sp) (dfs.h:101-253):
//   level-synchronous BFS reachability from sp on the device (the set of
//   reachable nodes does not depend on visit order), then the isolated
//   nodes in ascending id order, one wavefront applying the four repair
//   steps of try_repair_isolated_node (dfs.h:62-98) in sequence, as the
//   reference does, since each repair can fill a row a later one would use.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "internal.h"

namespace nbpg {

// ---------------------------------------------------------------------------
// BFS reachability
// ---------------------------------------------------------------------------

// one thread per (frontier node, adjacency slot); newly reached nodes are
// appended to `next` with one atomic per wavefront
__global__ __launch_bounds__(256) void bfs_level_kernel(const uint32_t *__restrict__ nbr,
                                                        uint32_t R, const uint32_t *__restrict__ front,
                                                        uint64_t nf, uint32_t *__restrict__ flag,
                                                        uint32_t *__restrict__ next,
                                                        uint32_t *__restrict__ next_cnt)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x % WAVE;
    bool fresh = false;
    uint32_t u = NBPG_SENT;
    if (t < nf * R) {
        const uint32_t v = front[t / R];
        u = nbr[(uint64_t)v * R + (t % R)];
        if (u != NBPG_SENT && flag[u] == 0)
            fresh = atomicCAS(&flag[u], 0u, 1u) == 0u;
    }
    const uint64_t m = ballot(fresh);
    if (m == 0)
        return;
    const int first = __builtin_ctzll(m);
    uint32_t base = 0;
    if ((int)lane == first)
        base = atomicAdd(next_cnt, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, first);
    if (fresh)
        next[base + prefix_count(m)] = u;
}

// every node left unreached, in no particular order
__global__ __launch_bounds__(256) void unreached_kernel(const uint32_t *__restrict__ flag, uint64_t n,
                                                        uint32_t *__restrict__ out,
                                                        uint32_t *__restrict__ cnt)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x % WAVE;
    const bool iso = i < n && flag[i] == 0;
    const uint64_t m = ballot(iso);
    if (m == 0)
        return;
    const int first = __builtin_ctzll(m);
    uint32_t base = 0;
    if ((int)lane == first)
        base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, first);
    if (iso)
        out[base + prefix_count(m)] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// repair: one wavefront, isolated nodes in ascending order
// ---------------------------------------------------------------------------

// rows are read and written at device scope: a row appended to by lane 0
// may be read back by any lane for a later node
__device__ __forceinline__ uint32_t ld_dev(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// add_edge_if_possible (dfs.h:45-58): a full row refuses; an existing edge
// counts as added; otherwise `to` is appended
__device__ bool add_edge(uint32_t *nbr, uint32_t *deg, uint32_t R, uint32_t from, uint32_t to,
                         uint32_t lane)
{
    const uint32_t df = ld_dev(&deg[from]);
    if (df >= R)
        return false;
    bool has = false;
    for (uint32_t k = lane; k < df; k += WAVE)
        has |= ld_dev(&nbr[(uint64_t)from * R + k]) == to;
    if (ballot(has) != 0)
        return true;
    if (lane == 0) {
        st_dev(&nbr[(uint64_t)from * R + df], to);
        st_dev(&deg[from], df + 1);
    }
    __threadfence();
    return true;
}

__global__ __launch_bounds__(64) void repair_kernel(uint32_t *nbr, uint32_t *deg,
                                                    const uint32_t *__restrict__ flag, uint32_t R,
                                                    uint64_t n, const uint32_t *__restrict__ iso,
                                                    uint64_t niso, uint32_t sp)
{
    const uint32_t lane = threadIdx.x;
    // step 4 takes the first reachable row with room; rows only fill up, so
    // that row never moves backwards and the scan resumes where it stopped
    uint64_t cursor = 0;
    for (uint64_t t = 0; t < niso; ++t) {
        const uint32_t i = iso[t];
        const uint32_t d = ld_dev(&deg[i]);
        bool done = false;
        // 1. a neighbour in the reachable part takes the reverse edge
        for (uint32_t j = 0; j < d && !done; ++j) {
            const uint32_t nid = ld_dev(&nbr[(uint64_t)i * R + j]);
            if (flag[nid] == 1)
                done = add_edge(nbr, deg, R, nid, i, lane);
        }
        // 2. any neighbour with room
        for (uint32_t j = 0; j < d && !done; ++j)
            done = add_edge(nbr, deg, R, ld_dev(&nbr[(uint64_t)i * R + j]), i, lane);
        // 3. the entry
        if (!done && flag[sp] == 1)
            done = add_edge(nbr, deg, R, sp, i, lane);
        // 4. the first reachable row with room (a reachable row cannot
        //    already point at an unreachable node)
        while (!done && cursor < n) {
            const uint64_t r = cursor + lane;
            const bool ok = r < n && flag[r] == 1 && ld_dev(&deg[r]) < R;
            const uint64_t m = ballot(ok);
            if (m) {
                cursor += (uint64_t)__builtin_ctzll(m);
                done = add_edge(nbr, deg, R, (uint32_t)cursor, i, lane);
            } else {
                cursor += WAVE;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// pack / unpack of the file's id stream, nodes [a, b)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void pack_kernel(const uint32_t *__restrict__ nbr,
                                                   const uint32_t *__restrict__ deg,
                                                   const uint64_t *__restrict__ off, uint64_t a,
                                                   uint64_t cnt, uint32_t R,
                                                   uint32_t *__restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * R)
        return;
    const uint64_t v = a + t / R;
    const uint32_t j = (uint32_t)(t % R);
    if (j < deg[v])
        out[off[v] - off[a] + j] = nbr[v * R + j];
}

__global__ __launch_bounds__(256) void unpack_kernel(const uint32_t *__restrict__ in,
                                                     const uint32_t *__restrict__ fdeg,
                                                     const uint64_t *__restrict__ off, uint64_t a,
                                                     uint64_t cnt, uint32_t R, uint64_t n,
                                                     uint32_t *__restrict__ nbr,
                                                     uint32_t *__restrict__ deg,
                                                     uint32_t *__restrict__ bad)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * R)
        return;
    const uint64_t v = a + t / R;
    const uint32_t j = (uint32_t)(t % R);
    const uint32_t d = fdeg[v];
    uint32_t x = NBPG_SENT;
    if (j < d) {
        x = in[off[v] - off[a] + j];
        if (x >= n) {
            atomicOr(bad, 2u);
            x = NBPG_SENT;
        }
    }
    nbr[v * R + j] = x;
    if (j == 0)
        deg[v] = d;
}

// node ranges [a, b) of at most `cap` ids each (a node with more than cap
// ids gets a range of its own; cap >= R keeps that from happening)
static std::vector<uint64_t> ranges(const std::vector<uint64_t> &off, uint64_t n, uint64_t cap)
{
    std::vector<uint64_t> cut{0};
    uint64_t a = 0;
    while (a < n) {
        uint64_t b = a + 1;
        // the furthest b with off[b] - off[a] <= cap
        const uint64_t lim = off[a] + cap;
        b = (uint64_t)(std::upper_bound(off.begin() + a + 1, off.begin() + n + 1, lim) - off.begin()) - 1;
        if (b <= a)
            b = a + 1;
        cut.push_back(b);
        a = b;
    }
    return cut;
}

struct DevMem {
    void *p = nullptr;
    ~DevMem()
    {
        if (p)
            hipFree(p);
    }
    int alloc(size_t bytes)
    {
        if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) {
            p = nullptr;
            return set_error(NBPG_ENOMEM, "device buffer of %zu bytes", bytes);
        }
        return NBPG_OK;
    }
};

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f)
            fclose(f);
    }
};

static const uint64_t IO_CHUNK = 1ull << 26; // ids per pack / unpack step

int device_repair(nbpg_index *ix, uint32_t sp, uint64_t *out_isolated)
{
    const uint64_t n = ix->n;
    const uint32_t R = ix->R;
    if (sp >= n)
        return set_error(NBPG_EINVAL, "entry %u >= n", sp);
    DevMem flag, fa, fb, iso;
    NBPG_TRY(flag.alloc(n * 4));
    NBPG_TRY(fa.alloc(n * 4));
    NBPG_TRY(fb.alloc(n * 4));
    uint32_t *cnt = ix->counter + 64; // its own 128-byte line
    HIP_TRY(hipMemset(flag.p, 0, n * 4));
    const uint32_t one = 1;
    HIP_TRY(hipMemcpy((uint32_t *)flag.p + sp, &one, 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(fa.p, &sp, 4, hipMemcpyHostToDevice));
    uint32_t *front = (uint32_t *)fa.p, *next = (uint32_t *)fb.p;
    uint64_t nf = 1;
    while (nf) {
        HIP_TRY(hipMemset(cnt, 0, 4));
        const uint64_t th = nf * R;
        hipLaunchKernelGGL(bfs_level_kernel, dim3((uint32_t)((th + 255) / 256)), dim3(256), 0, 0,
                           ix->nbr, R, front, nf, (uint32_t *)flag.p, next, cnt);
        HIP_TRY(hipGetLastError());
        uint32_t h = 0;
        HIP_TRY(hipMemcpy(&h, cnt, 4, hipMemcpyDeviceToHost));
        nf = h;
        std::swap(front, next);
    }
    // the unreached nodes, sorted ascending on the host (a short list)
    HIP_TRY(hipMemset(cnt, 0, 4));
    hipLaunchKernelGGL(unreached_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0,
                       (const uint32_t *)flag.p, n, front, cnt);
    HIP_TRY(hipGetLastError());
    uint32_t niso = 0;
    HIP_TRY(hipMemcpy(&niso, cnt, 4, hipMemcpyDeviceToHost));
    if (niso) {
        std::vector<uint32_t> list(niso);
        HIP_TRY(hipMemcpy(list.data(), front, (size_t)niso * 4, hipMemcpyDeviceToHost));
        std::sort(list.begin(), list.end());
        HIP_TRY(hipMemcpy(front, list.data(), (size_t)niso * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(repair_kernel, dim3(1), dim3(64), 0, 0, ix->nbr, ix->deg,
                           (const uint32_t *)flag.p, R, n, front, (uint64_t)niso, sp);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    if (out_isolated)
        *out_isolated = niso;
    return NBPG_OK;
}

} // namespace nbpg

using namespace nbpg;

extern "C" {

int nbpg_index_repair(nbpg_index *ix, uint32_t sp, uint64_t *out_isolated)
{
    if (!ix)
        return set_error(NBPG_EINVAL, "null argument");
    return device_repair(ix, sp, out_isolated);
}

int nbpg_index_save_graph(const nbpg_index *ix, const char *path)
{
    if (!ix || !path)
        return set_error(NBPG_EINVAL, "null argument");
    const uint64_t n = ix->n;
    const uint32_t R = ix->R;
    if (n > 0xFFFFFFFFull)
        return set_error(NBPG_EUNSUP, "n does not fit the file's u32 header");
    std::vector<uint32_t> deg(n);
    HIP_TRY(hipMemcpy(deg.data(), ix->deg, n * 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> off(n + 1, 0);
    for (uint64_t i = 0; i < n; ++i)
        off[i + 1] = off[i] + deg[i];
    File f;
    f.f = fopen(path, "wb");
    if (!f.f)
        return set_error(NBPG_EINVAL, "cannot write %s", path);
    const uint32_t hdr[2] = {(uint32_t)n, R};
    if (fwrite(hdr, 4, 2, f.f) != 2 || fwrite(deg.data(), 4, n, f.f) != n)
        return set_error(NBPG_EINVAL, "write failed: %s", path);
    const uint64_t cap = std::max<uint64_t>(IO_CHUNK, R);
    DevMem doff, dbuf;
    NBPG_TRY(doff.alloc((n + 1) * 8));
    NBPG_TRY(dbuf.alloc(std::min<uint64_t>(cap, std::max<uint64_t>(off[n], 1)) * 4));
    HIP_TRY(hipMemcpy(doff.p, off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
    std::vector<uint32_t> host(std::min<uint64_t>(cap, std::max<uint64_t>(off[n], 1)));
    const std::vector<uint64_t> cut = ranges(off, n, cap);
    for (size_t r = 0; r + 1 < cut.size(); ++r) {
        const uint64_t a = cut[r], b = cut[r + 1], m = off[b] - off[a];
        if (m == 0)
            continue;
        const uint64_t th = (b - a) * R;
        hipLaunchKernelGGL(pack_kernel, dim3((uint32_t)((th + 255) / 256)), dim3(256), 0, 0, ix->nbr,
                           ix->deg, (const uint64_t *)doff.p, a, b - a, R, (uint32_t *)dbuf.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpy(host.data(), dbuf.p, m * 4, hipMemcpyDeviceToHost));
        if (fwrite(host.data(), 4, m, f.f) != m)
            return set_error(NBPG_EINVAL, "write failed: %s", path);
    }
    return NBPG_OK;
}

int nbpg_index_load_graph(nbpg_index *ix, const char *path)
{
    if (!ix || !path)
        return set_error(NBPG_EINVAL, "null argument");
    File f;
    f.f = fopen(path, "rb");
    if (!f.f)
        return set_error(NBPG_EINVAL, "cannot open %s", path);
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f.f) != 2)
        return set_error(NBPG_EINVAL, "short file %s", path);
    const uint64_t n = hdr[0];
    const uint32_t fR = hdr[1];
    if (n != ix->n || fR > ix->R)
        return set_error(NBPG_EINVAL, "%s holds %llu nodes of degree <= %u; the index has %llu, R %u",
                         path, (unsigned long long)n, fR, (unsigned long long)ix->n, ix->R);
    std::vector<uint32_t> deg(n);
    if (fread(deg.data(), 4, n, f.f) != n)
        return set_error(NBPG_EINVAL, "short file %s", path);
    std::vector<uint64_t> off(n + 1, 0);
    for (uint64_t i = 0; i < n; ++i) {
        if (deg[i] > fR)
            return set_error(NBPG_EINVAL, "%s: degree %u > R %u at node %llu", path, deg[i], fR,
                             (unsigned long long)i);
        off[i + 1] = off[i] + deg[i];
    }
    const uint32_t R = ix->R;
    const uint64_t cap = std::max<uint64_t>(IO_CHUNK, R);
    DevMem doff, ddeg, dbuf;
    NBPG_TRY(doff.alloc((n + 1) * 8));
    NBPG_TRY(ddeg.alloc(n * 4));
    NBPG_TRY(dbuf.alloc(std::min<uint64_t>(cap, std::max<uint64_t>(off[n], 1)) * 4));
    HIP_TRY(hipMemcpy(doff.p, off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ddeg.p, deg.data(), n * 4, hipMemcpyHostToDevice));
    uint32_t *bad = ix->counter + 32;
    HIP_TRY(hipMemset(bad, 0, 4));
    std::vector<uint32_t> host(std::min<uint64_t>(cap, std::max<uint64_t>(off[n], 1)));
    const std::vector<uint64_t> cut = ranges(off, n, cap);
    for (size_t r = 0; r + 1 < cut.size(); ++r) {
        const uint64_t a = cut[r], b = cut[r + 1], m = off[b] - off[a];
        if (m && fread(host.data(), 4, m, f.f) != m)
            return set_error(NBPG_EINVAL, "short file %s", path);
        if (m)
            HIP_TRY(hipMemcpy(dbuf.p, host.data(), m * 4, hipMemcpyHostToDevice));
        const uint64_t th = (b - a) * R;
        hipLaunchKernelGGL(unpack_kernel, dim3((uint32_t)((th + 255) / 256)), dim3(256), 0, 0,
                           (const uint32_t *)dbuf.p, (const uint32_t *)ddeg.p,
                           (const uint64_t *)doff.p, a, b - a, R, n, ix->nbr, ix->deg, bad);
        HIP_TRY(hipGetLastError());
    }
    uint32_t hbad = 0;
    HIP_TRY(hipMemcpy(&hbad, bad, 4, hipMemcpyDeviceToHost));
    if (hbad)
        return set_error(NBPG_EINVAL, "%s: neighbour id >= n", path);
    return NBPG_OK;
}

} // extern "C"
