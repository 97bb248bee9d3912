"""nbpg_amd -- MI355X-native NBPG hot path (beam search, RobustPrune, CompGT).

Python front end over the C-ABI in include/nbpg.h.  The compute lives in
nbpg_amd/libnbpg_amd.so (hand-written gfx950 HIP); this module only moves
buffers: device memory comes from torch tensors on the current HIP device.
There is no CPU fallback: every entry point raises when the library is
missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbpg_amd.so")

U8, I8, U4X2, I4X2, F32 = 0, 1, 2, 3, 4
PCA_PLAIN, PCA_BF16 = 0, 1
CODE_F32, CODE_U8, CODE_I8, CODE_U4, CODE_I4 = 0, 1, 2, 3, 4
# PCACompress picks the code type from the output extension (PCACompress.cpp:25-37)
CODE_OF_EXT = {".fbin": CODE_F32, ".u8bin": CODE_U8, ".i8bin": CODE_I8, ".u4bin": CODE_U4,
               ".i4bin": CODE_I4}
SEARCH_REF, SEARCH_FAST, SEARCH_FRONTIER, SEARCH_EDGES, SEARCH_VISITED = 0, 1, 0x10, 0x20, 0x40
SEARCH_SEED_NEAREST = 0x80  # FAST: each query starts from its nearest seed row (set_seeds)
MAX_SEEDS = 4096


def visited_cap(L):
    """ids per query of a visited list (NBPG_VISITED_CAP): 8L, at most 1024"""
    return min(8 * L, 1024)
METRIC_L2, METRIC_MIPS = 0, 1
GT_PLAIN, GT_MFMA, GT_REFTIES, GT_MIPS = 0, 1, 0x100, 0x200
SENT = 0xFFFFFFFF
DTYPE_OF_EXT = {".u8bin": U8, ".i8bin": I8, ".u4bin": U4X2, ".i4bin": I4X2}

_lib = None


class NbpgError(RuntimeError):
    pass


class QueryParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("cut", C.c_double), ("L", C.c_uint32),
                ("limit", C.c_uint64), ("degree_limit", C.c_uint32)]


class FastConfig(C.Structure):
    _fields_ = [("waves_per_group", C.c_uint32), ("expand_width", C.c_uint32),
                ("filter_log2", C.c_uint32), ("groups", C.c_uint32), ("row_split", C.c_uint32)]


class BuildParams(C.Structure):
    _fields_ = [("R", C.c_uint32), ("L", C.c_uint32), ("VL", C.c_uint32),
                ("top_degree", C.c_uint32), ("num_split", C.c_uint32), ("num_passes", C.c_uint32),
                ("limit", C.c_uint64), ("min_alpha", C.c_double), ("max_alpha", C.c_double),
                ("min_cos_angle", C.c_double), ("max_cos_angle", C.c_double),
                ("dynamic_prune", C.c_int32), ("rand_sp", C.c_int32),
                ("repair_isolate", C.c_int32), ("sp_seed", C.c_uint64)]


VP = C.c_void_p
U32, U64, I32, DBL = C.c_uint32, C.c_uint64, C.c_int, C.c_double


class ServerConfig(C.Structure):
    _fields_ = [("slots", U32), ("max_batch_queries", U32), ("mem_budget", U64),
                ("max_L", U32), ("max_k", U32), ("fast", FastConfig), ("sp_num", U32),
                ("start_points", U32 * 64), ("seed_nearest", U32)]


class ServerStats(C.Structure):
    _fields_ = [("requests", U64), ("queries", U64), ("batches", U64),
                ("mean_batch_queries", DBL), ("p50_ms", DBL), ("p99_ms", DBL), ("max_ms", DBL),
                ("batch_capacity", U32)]

_SIGS = {
    "nbpg_index_save_graph": (I32, [VP, C.c_char_p]),
    "nbpg_index_load_graph": (I32, [VP, C.c_char_p]),
    "nbpg_index_repair": (I32, [VP, U32, C.POINTER(U64)]),
    "nbpg_last_error": (C.c_char_p, []),
    "nbpg_init": (I32, [I32]),
    "nbpg_device_info": (I32, [C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "nbpg_index_create": (I32, [C.POINTER(VP), I32, U64, U32, U32]),
    "nbpg_index_destroy": (I32, [VP]),
    "nbpg_index_set_points": (I32, [VP, VP, U64, U64, U64]),
    "nbpg_index_set_graph": (I32, [VP, VP, U64, U64]),
    "nbpg_index_get_graph": (I32, [VP, VP, U64, U64]),
    "nbpg_index_info": (I32, [VP, C.POINTER(I32), C.POINTER(U64), C.POINTER(U32), C.POINTER(U32),
                              C.POINTER(U32)]),
    "nbpg_l2_batch": (I32, [VP, VP, VP, U64, VP, VP]),
    "nbpg_l2_pairs": (I32, [I32, VP, VP, U64, U32, VP, VP]),
    "nbpg_search_rerank": (I32, [VP, VP, U64, U64, I32, VP, U32, VP, U64, VP, U32,
                                 C.POINTER(QueryParams), U32, I32, VP, VP, VP, VP, VP]),
    "nbpg_search_rerank_noalloc": (I32, [VP, VP, U64, U64, I32, VP, U32, VP, U64, VP, U32,
                                         C.POINTER(QueryParams), U32, I32, VP, VP, VP, VP, VP]),
    "nbpg_search_graph_create": (I32, [C.POINTER(VP), VP, VP, U64, U64, VP, U32,
                                       C.POINTER(QueryParams), I32, C.POINTER(FastConfig),
                                       VP, VP, VP]),
    "nbpg_search_graph_launch": (I32, [VP, VP]),
    "nbpg_search_uneven_graph_create": (I32, [C.POINTER(VP), VP, VP, U64, U64, VP, VP, U32, U32,
                                              VP, U32, C.POINTER(FastConfig), VP, VP, VP]),
    "nbpg_pick_seeds": (I32, [VP, U32, U32, U32, U64, VP, C.POINTER(U32)]),
    "nbpg_search_graph_destroy": (I32, [VP]),
    "nbpg_search": (I32, [VP, VP, U64, U64, VP, U32, C.POINTER(QueryParams), I32,
                          C.POINTER(FastConfig), VP, VP, VP, VP]),
    "nbpg_search_uneven": (I32, [VP, VP, U64, U64, VP, VP, U32, U32, VP, U32,
                                 C.POINTER(FastConfig), VP, VP, VP, VP]),
    "nbpg_robust_prune_batch": (I32, [VP, VP, U64, VP, VP, U32, VP, DBL, DBL, VP, VP, VP, VP]),
    "nbpg_compgt": (I32, [I32, VP, U64, VP, U64, U32, U32, I32, VP, VP, VP]),
    "nbpg_build": (I32, [VP, C.POINTER(BuildParams), C.POINTER(U64)]),
    "nbpg_server_create": (I32, [C.POINTER(VP), VP, C.POINTER(ServerConfig)]),
    "nbpg_server_submit": (I32, [VP, VP, U32, U64, U32, U32, C.POINTER(U64)]),
    "nbpg_server_wait": (I32, [VP, U64, VP, VP, C.POINTER(DBL)]),
    "nbpg_server_stats": (I32, [VP, I32, C.POINTER(ServerStats)]),
    "nbpg_server_submit_many": (I32, [VP, VP, U32, VP, U64, U32, U32, VP]),
    "nbpg_server_wait_many": (I32, [VP, U32, VP, VP, VP, VP]),
    "nbpg_server_destroy": (I32, [VP]),
    "nbpg_rerank": (I32, [I32, I32, VP, U64, U32, VP, U64, VP, U32, U32, U32, U32, VP, VP, VP]),
    "nbpg_index_set_metric": (I32, [VP, I32]),
    "nbpg_index_set_seeds": (I32, [VP, VP, U32]),
    "nbpg_pca_project": (I32, [I32, VP, U64, U32, VP, VP, U32, I32, I32, VP, VP, VP]),
    "nbpg_pca_quantize": (I32, [VP, U64, U32, I32, VP, VP]),
    "nbpg_pca_stats_f64": (I32, [I32, VP, U64, U32, VP, VP, VP]),
}


def lib():
    """Load libnbpg_amd.so; raises NbpgError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NbpgError("libnbpg_amd.so is not built (run __graft_entry__.build() or python nbpg_amd/make_lib.py)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().nbpg_last_error().decode(errors="replace")
        raise NbpgError("%s failed (%d): %s" % (what, rc, msg))


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise NbpgError("no HIP device visible")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def init(device=0):
    _check(lib().nbpg_init(device), "nbpg_init")


def device_info():
    cu, lds, l2 = I32(), I32(), I32()
    _check(lib().nbpg_device_info(C.byref(cu), C.byref(lds), C.byref(l2)), "nbpg_device_info")
    return {"cu": cu.value, "lds_per_group": lds.value, "l2": l2.value}


class Index:
    """Device-resident vectors + adjacency (see include/nbpg.h)."""

    def __init__(self, dtype, n, dims, R):
        _torch()
        h = VP()
        _check(lib().nbpg_index_create(C.byref(h), dtype, n, dims, R), "nbpg_index_create")
        self.h, self.dtype, self.n, self.dims, self.R = h, dtype, n, dims, R

    def close(self):
        if self.h:
            lib().nbpg_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, rows, first=0):
        a = np.ascontiguousarray(rows)
        assert a.ndim == 2 and a.shape[1] * a.itemsize >= self.dims
        _check(lib().nbpg_index_set_points(self.h, a.ctypes.data, a.shape[1] * a.itemsize, first,
                                           a.shape[0]), "nbpg_index_set_points")

    def set_seeds(self, ids):
        """seed rows for SEARCH_SEED_NEAREST (distinct ids < n, at most
        MAX_SEEDS; an empty list drops them)"""
        a = np.ascontiguousarray(ids, dtype=np.uint32)
        _check(lib().nbpg_index_set_seeds(self.h, a.ctypes.data, len(a)), "nbpg_index_set_seeds")

    def set_graph(self, rows, first=0):
        a = np.ascontiguousarray(rows, dtype=np.uint32)
        assert a.ndim == 2 and a.shape[1] == self.R + 1
        _check(lib().nbpg_index_set_graph(self.h, a.ctypes.data, first, a.shape[0]),
               "nbpg_index_set_graph")

    def get_graph(self, first=0, count=None):
        count = self.n - first if count is None else count
        a = np.zeros((count, self.R + 1), np.uint32)
        _check(lib().nbpg_index_get_graph(self.h, a.ctypes.data, first, count),
               "nbpg_index_get_graph")
        return a

    def set_metric(self, metric):
        """METRIC_L2 (the default) or METRIC_MIPS for every call on this index."""
        _check(lib().nbpg_index_set_metric(self.h, metric), "nbpg_index_set_metric")

    def save_graph(self, path):
        """Write the resident graph as a reference .graph file."""
        _check(lib().nbpg_index_save_graph(self.h, os.fsencode(path)), "nbpg_index_save_graph")

    def load_graph(self, path):
        """Read a reference .graph file into the resident graph."""
        _check(lib().nbpg_index_load_graph(self.h, os.fsencode(path)), "nbpg_index_load_graph")

    def repair(self, sp=0):
        """Isolated-node repair from entry sp; returns how many nodes were
        unreachable."""
        c = U64()
        _check(lib().nbpg_index_repair(self.h, sp, C.byref(c)), "nbpg_index_repair")
        return c.value

    @property
    def row_stride(self):
        s = U32()
        _check(lib().nbpg_index_info(self.h, None, None, None, None, C.byref(s)), "nbpg_index_info")
        return s.value


def to_device(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def l2_batch(index, q_dev, ids_dev):
    torch = _torch()
    out = torch.empty(ids_dev.numel(), dtype=torch.float32, device="cuda")
    _check(lib().nbpg_l2_batch(index.h, _ptr(q_dev), _ptr(ids_dev), ids_dev.numel(), _ptr(out),
                               _stream()), "nbpg_l2_batch")
    return out


def search(index, q_dev, L, k, mode=SEARCH_REF, cut=1.35, limit=None, degree_limit=None,
           entry=0, start_points=None, fast=None, stats=True, out=None, frontier=False,
           edges=False, visited=False):
    """Beam search for every row of q_dev (uint8 tensor, nq x stride bytes).
    frontier=True returns each query's whole frontier (nq x L) instead of
    its first k (the reference mode still cuts with k).  edges=True (FAST
    mode): stats rows are {visited, dist_cmps, adjacency entries read}.
    visited=True (FAST mode): ids are each query's expanded nodes in
    expansion order (nq x visited_cap(L), 0xFFFFFFFF past the count)."""
    torch = _torch()
    nq = q_dev.shape[0]
    qstride = q_dev.shape[1] * q_dev.element_size()
    qp = QueryParams(k, cut, L, index.n if limit is None else limit,
                     index.R if degree_limit is None else degree_limit)
    sp = np.array([entry] if start_points is None else start_points, dtype=np.uint32)
    width = visited_cap(L) if visited else (L if frontier else k)
    if frontier:
        mode |= SEARCH_FRONTIER
    if edges:
        mode |= SEARCH_EDGES
    if visited:
        mode |= SEARCH_VISITED
    if out is None:
        ids = torch.empty((nq, width), dtype=torch.int32, device="cuda")
        d = torch.empty((nq, width), dtype=torch.float32, device="cuda")
        st = torch.empty((nq, 3 if edges else 2), dtype=torch.int32, device="cuda") \
            if stats else None
    else:
        ids, d, st = out
    fc = FastConfig(*fast) if fast is not None else None
    _check(lib().nbpg_search(index.h, _ptr(q_dev), nq, qstride, sp.ctypes.data, len(sp),
                             C.byref(qp), mode, C.byref(fc) if fc is not None else None,
                             _ptr(ids), _ptr(d), _ptr(st), _stream()), "nbpg_search")
    return ids, d, st


class SearchGraph:
    """One search captured as a HIP graph (nbpg_search_graph_*): replay()
    searches whatever q_dev holds into the same outputs (ids, dists, stats)."""

    def __init__(self, index, q_dev, L, k, mode=SEARCH_FAST, cut=1.35, limit=None,
                 degree_limit=None, entry=0, start_points=None, fast=None, edges=False):
        torch = _torch()
        nq = q_dev.shape[0]
        qstride = q_dev.shape[1] * q_dev.element_size()
        qp = QueryParams(k, cut, L, index.n if limit is None else limit,
                         index.R if degree_limit is None else degree_limit)
        sp = np.array([entry] if start_points is None else start_points, dtype=np.uint32)
        if edges:
            mode |= SEARCH_EDGES
        self.ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        self.dists = torch.empty((nq, k), dtype=torch.float32, device="cuda")
        self.stats = torch.empty((nq, 3 if edges else 2), dtype=torch.int32, device="cuda")
        fc = FastConfig(*fast) if fast is not None else None
        h = VP()
        _check(lib().nbpg_search_graph_create(C.byref(h), index.h, _ptr(q_dev), nq, qstride,
                                              sp.ctypes.data, len(sp), C.byref(qp), mode,
                                              C.byref(fc) if fc is not None else None,
                                              _ptr(self.ids), _ptr(self.dists),
                                              _ptr(self.stats)), "nbpg_search_graph_create")
        self.h, self.index, self.q_dev = h, index, q_dev

    def replay(self):
        _check(lib().nbpg_search_graph_launch(self.h, _stream()), "nbpg_search_graph_launch")
        return self.ids, self.dists, self.stats

    def close(self):
        if getattr(self, "h", None):
            lib().nbpg_search_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search_rerank(index, q_dev, raw_base_dev, raw_q_dev, L, k, rerank_factor=100,
                  mode=SEARCH_REF, cut=1.35, limit=None, degree_limit=None, entry=0,
                  start_points=None, fast=None, raw_dtype=None, noalloc=False):
    """Search over the index's quantized rows with full-precision rows for the
    final distances (beam_search_rerank): rerank_factor > 0 re-scores the
    frontier head, 0 is the filtering mode (SEARCH_REF only).  raw_base_dev:
    n x raw_dims (float32 / uint8 / int8 tensor), raw_q_dev: nq x raw_dims.
    raw_dtype overrides the type taken from the tensor (U4X2 / I4X2 rows).
    noalloc: call nbpg_search_rerank_noalloc (scratch in the library's fixed
    per-GPU array, no device allocation by the library).
    Returns (ids, dists, stats), nq x k."""
    torch = _torch()
    nq = q_dev.shape[0]
    qstride = q_dev.shape[1] * q_dev.element_size()
    qp = QueryParams(k, cut, L, index.n if limit is None else limit,
                     index.R if degree_limit is None else degree_limit)
    sp = np.array([entry] if start_points is None else start_points, dtype=np.uint32)
    ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    d = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    st = torch.empty((nq, 2), dtype=torch.int32, device="cuda")
    fc = FastConfig(*fast) if fast is not None else None
    raw_dims = raw_base_dev.shape[1]
    rdt = _x_dtype(raw_base_dev) if raw_dtype is None else raw_dtype
    name = "nbpg_search_rerank_noalloc" if noalloc else "nbpg_search_rerank"
    _check(getattr(lib(), name)(index.h, _ptr(q_dev), nq, qstride, rdt, _ptr(raw_base_dev),
                                raw_dims, _ptr(raw_q_dev),
                                raw_q_dev.shape[1] * raw_q_dev.element_size(), sp.ctypes.data,
                                len(sp), C.byref(qp), rerank_factor, mode,
                                C.byref(fc) if fc is not None else None, _ptr(ids), _ptr(d),
                                _ptr(st), _stream()), name)
    return ids, d, st


class UnevenSearchGraph(SearchGraph):
    """nbpg_search_uneven captured as a HIP graph (nbpg_search_uneven_graph_create):
    replay() searches whatever q_dev, L_dev and k_dev hold into the same
    outputs."""

    def __init__(self, index, q_dev, L_dev, k_dev, max_L, out_k, entry=0, fast=None,
                 start_points=None):
        torch = _torch()
        nq = q_dev.shape[0]
        qstride = q_dev.shape[1] * q_dev.element_size()
        sp = np.array([entry] if start_points is None else start_points, dtype=np.uint32)
        self.ids = torch.empty((nq, out_k), dtype=torch.int32, device="cuda")
        self.dists = torch.empty((nq, out_k), dtype=torch.float32, device="cuda")
        self.stats = torch.empty((nq, 2), dtype=torch.int32, device="cuda")
        fc = FastConfig(*fast) if fast is not None else None
        h = VP()
        _check(lib().nbpg_search_uneven_graph_create(
            C.byref(h), index.h, _ptr(q_dev), nq, qstride, _ptr(L_dev), _ptr(k_dev), max_L, out_k,
            sp.ctypes.data, len(sp), C.byref(fc) if fc is not None else None, _ptr(self.ids),
            _ptr(self.dists), _ptr(self.stats)), "nbpg_search_uneven_graph_create")
        self.h, self.index, self.q_dev = h, index, q_dev


def search_uneven(index, q_dev, L_dev, k_dev, max_L, out_k, entry=0, fast=None, out=None,
                  start_points=None):
    """FAST search with per-query L and k (device u32 arrays); start points:
    start_points (up to 64 ids) or the single entry."""
    torch = _torch()
    sp = np.array([entry] if start_points is None else start_points, dtype=np.uint32)
    nq = q_dev.shape[0]
    qstride = q_dev.shape[1] * q_dev.element_size()
    if out is None:
        ids = torch.empty((nq, out_k), dtype=torch.int32, device="cuda")
        d = torch.empty((nq, out_k), dtype=torch.float32, device="cuda")
        st = torch.empty((nq, 2), dtype=torch.int32, device="cuda")
    else:
        ids, d, st = out
    fc = FastConfig(*fast) if fast is not None else None
    _check(lib().nbpg_search_uneven(index.h, _ptr(q_dev), nq, qstride, _ptr(L_dev), _ptr(k_dev),
                                    max_L, out_k, sp.ctypes.data, len(sp),
                                    C.byref(fc) if fc is not None else None,
                                    _ptr(ids), _ptr(d), _ptr(st), _stream()), "nbpg_search_uneven")
    return ids, d, st


class Server:
    """Continuous-batching request loop over a resident index (see
    include/nbpg.h).  submit() and wait() may be called from any thread;
    wait() releases the GIL while it blocks."""

    def __init__(self, index, slots=3, max_batch_queries=4096, mem_budget=0, max_L=256,
                 max_k=100, entry=0, fast=(2, 1, 11, 0, 0), start_points=None,
                 seed_nearest=False):
        _torch()
        self.index = index
        sp = [entry] if start_points is None else [int(x) for x in start_points]
        if not 1 <= len(sp) <= 64:
            raise ValueError("1..64 start points")
        cfg = ServerConfig(slots, max_batch_queries, mem_budget, max_L, max_k, FastConfig(*fast),
                           len(sp), (U32 * 64)(*sp), int(bool(seed_nearest)))
        h = VP()
        _check(lib().nbpg_server_create(C.byref(h), index.h, C.byref(cfg)), "nbpg_server_create")
        self.h = h
        self._shape = {}

    def submit(self, rows, L, k):
        a = np.ascontiguousarray(rows)
        a = a.view(np.uint8).reshape(a.shape[0], -1)
        t = U64()
        _check(lib().nbpg_server_submit(self.h, a.ctypes.data, a.shape[0], a.shape[1], L, k,
                                        C.byref(t)), "nbpg_server_submit")
        self._shape[t.value] = (a.shape[0], k)
        return t.value

    def wait(self, ticket):
        nq, k = self._shape.pop(ticket)
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        lat = DBL()
        _check(lib().nbpg_server_wait(self.h, ticket, ids.ctypes.data, d.ctypes.data,
                                      C.byref(lat)), "nbpg_server_wait")
        return ids, d, lat.value

    def submit_many(self, rows, sizes, L, k):
        """requests of sizes[i] consecutive rows each (rows holds sum(sizes));
        returns the tickets (uint64 array)"""
        a = np.ascontiguousarray(rows)
        a = a.view(np.uint8).reshape(a.shape[0], -1)
        sz = np.ascontiguousarray(sizes, dtype=np.uint32)
        if int(sz.sum()) != a.shape[0]:
            raise ValueError("sizes must add up to the rows given")
        t = np.empty(len(sz), np.uint64)
        _check(lib().nbpg_server_submit_many(self.h, a.ctypes.data, len(sz), sz.ctypes.data,
                                             a.shape[1], L, k, t.ctypes.data),
               "nbpg_server_submit_many")
        for tk, m in zip(t.tolist(), sz.tolist()):
            self._shape[tk] = (m, k)
        return t

    def wait_many(self, tickets):
        """(ids, dists, latencies_ms) of the tickets' requests, rows in ticket
        order"""
        t = np.ascontiguousarray(tickets, dtype=np.uint64)
        shapes = [self._shape.pop(tk) for tk in t.tolist()]
        total = sum(m * k for m, k in shapes)
        ids = np.empty(total, np.uint32)
        d = np.empty(total, np.float32)
        lat = np.empty(len(t), np.float64)
        _check(lib().nbpg_server_wait_many(self.h, len(t), t.ctypes.data, ids.ctypes.data,
                                           d.ctypes.data, lat.ctypes.data),
               "nbpg_server_wait_many")
        return ids, d, lat

    def stats(self, reset=False):
        st = ServerStats()
        _check(lib().nbpg_server_stats(self.h, int(reset), C.byref(st)), "nbpg_server_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().nbpg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def robust_prune_batch(index, nodes_dev, cand_dev, offsets_dev, R, alpha=1.2, cos_angle=0.0,
                       alphas_dev=None):
    """cand_dev: int64 tensor of packed (id u32, dist f32) records."""
    torch = _torch()
    n = nodes_dev.numel()
    ids = torch.empty((n, R), dtype=torch.int32, device="cuda")
    d = torch.empty((n, R), dtype=torch.float32, device="cuda")
    cnt = torch.empty(n, dtype=torch.int32, device="cuda")
    _check(lib().nbpg_robust_prune_batch(index.h, _ptr(nodes_dev), n, _ptr(cand_dev),
                                         _ptr(offsets_dev), R, _ptr(alphas_dev), alpha, cos_angle,
                                         _ptr(ids), _ptr(d), _ptr(cnt), _stream()),
           "nbpg_robust_prune_batch")
    return ids, d, cnt


def compgt(base_dev, q_dev, k, dtype=U8, mode=GT_PLAIN):
    torch = _torch()
    nq = q_dev.shape[0]
    ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    d = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    _check(lib().nbpg_compgt(dtype, _ptr(base_dev), base_dev.shape[0], _ptr(q_dev), nq,
                             base_dev.shape[1], k, mode, _ptr(ids), _ptr(d), _stream()),
           "nbpg_compgt")
    return ids, d


def rerank(raw_base_dev, raw_q_dev, cand_dev, k, rerank_factor=100, raw_dtype=None,
           metric=None):
    """Re-score each query's frontier (cand_dev: nq x C ids, e.g. the ids of a
    search over codes with k = L) with the full-precision rows; returns
    (ids, dists) nq x k."""
    torch = _torch()
    nq, C = cand_dev.shape
    n, dims = raw_base_dev.shape
    if cand_dev.stride(1) != 1 or not raw_base_dev.is_contiguous() or not raw_q_dev.is_contiguous():
        raise NbpgError("rerank needs packed raw rows and unit-stride candidate rows")
    ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    d = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    rdt = _x_dtype(raw_base_dev) if raw_dtype is None else raw_dtype
    _check(lib().nbpg_rerank(rdt, METRIC_L2 if metric is None else metric, _ptr(raw_base_dev), n,
                             dims, _ptr(raw_q_dev),
                             nq, _ptr(cand_dev), cand_dev.stride(0), C, k, rerank_factor,
                             _ptr(ids), _ptr(d),
                             _stream()), "nbpg_rerank")
    return ids, d


def _x_dtype(t):
    torch = _torch()
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.uint8:
        return U8
    if t.dtype == torch.int8:
        return I8
    raise NbpgError("rows must be float32, uint8 or int8, not %s" % t.dtype)


def pca_project(x_dev, W_dev, bias_dev, code=CODE_I8, mode=PCA_BF16, want_y=False):
    """PCACompress's projection + quantizer over device rows x (n x d_in).
    Returns (codes or None, y or None)."""
    torch = _torch()
    n, d_in = x_dev.shape
    d_out = W_dev.shape[0]
    chunk = 2 if code in (CODE_U4, CODE_I4) else 1
    codes = None if code == CODE_F32 else torch.empty((n, d_out // chunk), dtype=torch.uint8,
                                                      device=x_dev.device)
    y = torch.empty((n, d_out), dtype=torch.float32, device=x_dev.device) \
        if (want_y or code == CODE_F32) else None
    _check(lib().nbpg_pca_project(_x_dtype(x_dev), _ptr(x_dev), n, d_in, _ptr(W_dev),
                                  _ptr(bias_dev), d_out, mode, code, _ptr(codes), _ptr(y),
                                  _stream()), "nbpg_pca_project")
    return codes, y


def pca_quantize(y_dev, code):
    torch = _torch()
    n, d = y_dev.shape
    chunk = 2 if code in (CODE_U4, CODE_I4) else 1
    out = torch.empty((n, d // chunk), dtype=torch.uint8, device=y_dev.device)
    _check(lib().nbpg_pca_quantize(_ptr(y_dev), n, d, code, _ptr(out), _stream()),
           "nbpg_pca_quantize")
    return out


def pca_stats(x_dev, cov=True):
    """FP64 (mean, covariance) of device rows x."""
    torch = _torch()
    n, d = x_dev.shape
    mean = torch.empty(d, dtype=torch.float64, device=x_dev.device)
    c = torch.empty((d, d), dtype=torch.float64, device=x_dev.device) if cov else None
    _check(lib().nbpg_pca_stats_f64(_x_dtype(x_dev), _ptr(x_dev), n, d, _ptr(mean), _ptr(c),
                                    _stream()), "nbpg_pca_stats_f64")
    return mean, c


def pick_seeds(index, count=64, sample=65536, iters=10, seed=0):
    """FAST-mode start nodes spread over the index's rows (nbpg_pick_seeds:
    k-means with `count` centres on a sample, then the row nearest each
    centre).  Returns distinct uint32 ids, ascending (at most count)."""
    out = np.zeros(count, np.uint32)
    m = U32()
    _check(lib().nbpg_pick_seeds(index.h, count, sample, iters, seed, out.ctypes.data,
                                 C.byref(m)), "nbpg_pick_seeds")
    return out[:m.value].copy()


def pack_cands(ids, dists):
    """(id u32, dist f32) records as one int64 array (little-endian)."""
    rec = np.zeros(len(ids), dtype=[("id", np.uint32), ("dist", np.float32)])
    rec["id"] = ids
    rec["dist"] = dists
    return rec.view(np.int64)


def build(index, R, L, alpha=1.2, passes=1, num_split=0, VL=140, min_alpha=1.0,
          dynamic_prune=False, rand_sp=False, repair=True, top_degree=0, sp_seed=1, limit=0,
          min_cos_angle=0.0, max_cos_angle=0.0):
    bp = BuildParams(R, L, VL, top_degree, num_split, passes, limit, min_alpha, alpha,
                     min_cos_angle, max_cos_angle, int(dynamic_prune), int(rand_sp), int(repair),
                     sp_seed)
    iso = U64(0)
    _check(lib().nbpg_build(index.h, C.byref(bp), C.byref(iso)), "nbpg_build")
    return iso.value
