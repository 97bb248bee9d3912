// internal.h -- host-side structures shared by the C-ABI translation units.
#pragma once
#include <stdarg.h>
#include <string.h>
#include <stddef.h>

#include <type_traits>
#include <utility>

#include "common.h"

// Device-resident index.  Layout in HBM:
//   vecs  : n rows x stride bytes, stride = 16 * nch (zero padded), so one
//           row is nch dwordx4 loads and any row of <= 128 B sits in one
//           128-B line when the stride is 128
//   norms : n x u32, |row|^2 in the type's dot units
//   nbr   : n x R u32 neighbour ids, unused slots = 0xFFFFFFFF (no degree
//           read on the search path: a row is one contiguous R*4-B gather)
//   deg   : n x u32 degrees (build and graph download)
struct nbpg_index {
    int dtype;
    uint64_t n;
    uint32_t dims;
    uint32_t R;
    uint32_t nch;
    uint32_t stride;
    uint8_t *vecs;
    uint32_t *norms;
    uint32_t *nbr;
    uint32_t *deg;
    uint32_t mips; // 0: squared L2, 1: inner product (negated dot)
    void *ws; // grow-only workspace for search/prune scratch
    size_t ws_bytes;
    // 128 u32 of device flags and counters, each user on its own 128-byte
    // line: [32] graph upload check, [40] prune flag, [64] BFS / repair
    // counts; then, from [128], NBPG_WORK_COUNTERS work counters of the
    // persistent searches (64 u32 each: the hand-out count, and at +48 the
    // exit count)
    uint32_t *counter;
    uint32_t launch_seq; // next work counter (host side, atomic)
    // seed rows for nearest-seed FAST starts (start.hip): n_seeds ids
    // ascending, their rows as signed bytes (stride each, zero past dims) and
    // their seed terms for squared L2 and the inner product, padded to whole
    // 16-row tiles
    uint32_t n_seeds;
    uint32_t *seed_ids;
    uint8_t *seed_rows;
    int32_t *seed_self_l2, *seed_self_ip;
};

namespace nbpg {

int set_error(int code, const char *fmt, ...);

#define HIP_TRY(x)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return nbpg::set_error(NBPG_EHIP, "%s failed: %s", #x, hipGetErrorString(e_));        \
    } while (0)

#define NBPG_TRY(x)                                                                                \
    do {                                                                                           \
        int r_ = (x);                                                                              \
        if (r_ != NBPG_OK)                                                                         \
            return r_;                                                                             \
    } while (0)

struct DevInfo {
    int device = -1;
    int cu = 0;
    int lds_max = 0;
    int l2 = 0;
    bool inited = false;
};
extern DevInfo g_dev;

int ensure_init();
// nearest-seed FAST starts (start.hip): the index's seed rows, and the pass
// that writes each query's start id to d_out_ids[q * out_k]
bool start_select_supported(const nbpg_index *ix);
int start_select(const nbpg_index *ix, const void *d_queries, uint64_t nq, uint64_t qstride,
                 uint32_t *d_out_ids, uint32_t out_k, hipStream_t stream);
void drop_seeds(nbpg_index *ix);
int ensure_ws(nbpg_index *idx, size_t bytes);
// A persistent search's work counter (zero when idle; the launch's last wave
// resets it).  Launches take the index's counters in turn, so launches on
// different streams may run at the same time as long as fewer than
// NBPG_WORK_COUNTERS of them are in flight on the device at once.
#define NBPG_WORK_COUNTERS 64
uint32_t *work_counter(const nbpg_index *ix);
// DFS::repair_isolate_node on the resident graph (graph_io.hip)
int device_repair(nbpg_index *ix, uint32_t sp, uint64_t *out_isolated);
// rerank of candidate rows (rerank.hip); raw query rows at q_stride bytes
#define RERANK_MAX 1024
int rerank_launch(int raw_dtype, uint32_t mips, const void *d_raw_base, uint64_t n, uint32_t dims,
                  const void *d_raw_queries, uint64_t q_stride, uint64_t nq,
                  const uint32_t *d_cand, uint32_t cand_stride, uint32_t n_cand, uint32_t k,
                  uint32_t rerank_factor, uint32_t *d_out_ids, float *d_out_dists,
                  hipStream_t stream);
// 16-byte chunks per row for `bytes` of payload: one of 2, 4, 6, 8, 16
uint32_t nch_for_bytes(uint32_t bytes);
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// calls f(std::integral_constant<int, DT>, std::integral_constant<int, NCH>)
// for the (dtype, nch) pair; the two packed 4-bit types share kernels
template <typename F> int dispatch_dt_nch(int dtype, uint32_t nch, F &&f)
{
    using std::integral_constant;
#define NBPG_NCH_CASES(DTV)                                                                        \
    switch (nch) {                                                                                 \
    case 2: f(integral_constant<int, DTV>(), integral_constant<int, 2>()); return NBPG_OK;         \
    case 4: f(integral_constant<int, DTV>(), integral_constant<int, 4>()); return NBPG_OK;         \
    case 6: f(integral_constant<int, DTV>(), integral_constant<int, 6>()); return NBPG_OK;         \
    case 8: f(integral_constant<int, DTV>(), integral_constant<int, 8>()); return NBPG_OK;         \
    case 16: f(integral_constant<int, DTV>(), integral_constant<int, 16>()); return NBPG_OK;       \
    default: break;                                                                                \
    }
    switch (dtype) {
    case NBPG_U8: NBPG_NCH_CASES(NBPG_U8) break;
    case NBPG_I8: NBPG_NCH_CASES(NBPG_I8) break;
    case NBPG_U4X2:
    case NBPG_I4X2: NBPG_NCH_CASES(NBPG_U4X2) break;
    default: break;
    }
#undef NBPG_NCH_CASES
    return set_error(NBPG_EUNSUP, "unsupported dtype %d / row chunks %u", dtype, nch);
}

} // namespace nbpg
