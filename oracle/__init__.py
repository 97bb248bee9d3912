"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module: it is the checker for the HIP path in nbpg_amd, never a
fallback for it.  `liboracle.so` is the C restatement of the reference
(oracle/nbpg_oracle.c); `_ref/libnbpg_ref.so` and the `_ref/` binaries are
the reference's own code compiled by oracle/Makefile where the reference
tree exists.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
U8, I8, U4X2, I4X2, F32 = 0, 1, 2, 3, 4
MIPS = 0x100  # OR into a dtype: the inner-product metric

_lib = None
_ref = None


class IdDist(C.Structure):
    _fields_ = [("id", C.c_uint32), ("dist", C.c_float)]


class Points(C.Structure):
    _fields_ = [("data", C.c_void_p), ("stride", C.c_uint64), ("n", C.c_uint64),
                ("dims", C.c_uint32), ("dtype", C.c_int32)]


class Graph(C.Structure):
    _fields_ = [("rows", C.c_void_p), ("n", C.c_uint64), ("R", C.c_uint32)]


class QParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("cut", C.c_double), ("L", C.c_uint32),
                ("limit", C.c_uint64), ("degree_limit", C.c_uint32), ("max_visited", C.c_uint64)]


class BuildParams(C.Structure):
    _fields_ = [("R", C.c_uint32), ("L", C.c_uint32), ("VL", C.c_uint32),
                ("top_degree", C.c_uint32), ("num_split", C.c_uint32), ("num_passes", C.c_uint32),
                ("limit", C.c_uint64), ("min_alpha", C.c_double), ("max_alpha", C.c_double),
                ("min_cos_angle", C.c_double), ("max_cos_angle", C.c_double),
                ("dynamic_prune", C.c_int32), ("rand_sp", C.c_int32),
                ("repair_isolate", C.c_int32), ("sp_seed", C.c_uint64)]


def build(ref=False):
    """Compile liboracle.so (and the reference pieces when ref=True)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.nbpgo_distance.restype = C.c_float
        L.nbpgo_distance.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32]
        L.nbpgo_hash64_2.restype = C.c_uint64
        L.nbpgo_hash64_2.argtypes = [C.c_uint64]
        L.nbpgo_hash_filter_size.restype = C.c_uint32
        L.nbpgo_hash_filter_size.argtypes = [C.c_uint32]
        L.nbpgo_compgt.restype = C.c_int
        L.nbpgo_compgt.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                   C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.nbpgo_sort_candidates.restype = C.c_uint64
        L.nbpgo_sort_candidates.argtypes = [C.c_uint32, C.c_void_p, C.c_uint64]
        L.nbpgo_robust_prune.restype = C.c_uint64
        L.nbpgo_robust_prune.argtypes = [C.c_uint32, C.c_void_p, C.POINTER(C.c_uint64),
                                         C.POINTER(Points), C.c_uint32, C.c_double, C.c_double,
                                         C.c_uint64]
        L.nbpgo_search.restype = C.c_int
        L.nbpgo_search.argtypes = [C.POINTER(Graph), C.POINTER(Points), C.POINTER(Points),
                                   C.c_void_p, C.c_uint32, C.POINTER(QParams), C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int]
        L.nbpgo_search_exact.restype = C.c_int
        L.nbpgo_search_exact.argtypes = [C.POINTER(Graph), C.POINTER(Points), C.POINTER(Points),
                                         C.POINTER(Points), C.POINTER(Points), C.c_void_p,
                                         C.c_uint32, C.POINTER(QParams), C.c_uint32, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int]
        L.nbpgo_build.restype = C.c_int
        L.nbpgo_build.argtypes = [C.POINTER(Points), C.POINTER(Graph), C.POINTER(BuildParams),
                                  C.POINTER(C.c_uint64), C.c_int]
        L.nbpgo_shuffle_mt19937.restype = None
        L.nbpgo_shuffle_mt19937.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32]
        L.nbpgo_sp_stream.restype = C.c_uint64
        L.nbpgo_sp_stream.argtypes = [C.c_uint64, C.c_uint64]
        L.nbpgo_repair_isolated.restype = C.c_uint64
        L.nbpgo_repair_isolated.argtypes = [C.POINTER(Graph), C.c_uint32]
        L.nbpgo_matmul.restype = None
        L.nbpgo_matmul.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.c_uint32, C.c_void_p, C.c_int]
        L.nbpgo_compress.restype = None
        L.nbpgo_compress.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_uint32,
                                     C.c_void_p, C.c_int]
        L.nbpgo_recall.restype = C.c_double
        L.nbpgo_recall.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_uint32,
                                   C.POINTER(Points), C.POINTER(Points), C.c_uint32]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libnbpg_ref.so"))


def host_cpu():
    """(model name, x86-64 ISA level 'v4' / 'v3' / None) of this host, from
    /proc/cpuinfo."""
    model, flags, ids = "unknown", set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key = line.split(":", 1)[0].strip()
                if key == "model name" and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif key in ("cpu family", "model", "stepping") and key not in ids:
                    ids[key] = line.split(":", 1)[1].strip()
                elif key == "flags" and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    if ids:
        # a masked model name (virtual machines) still carries the family
        model += " (family %s, model %s, stepping %s)" % (
            ids.get("cpu family", "?"), ids.get("model", "?"), ids.get("stepping", "?"))
    v3 = {"avx2", "fma", "bmi1", "bmi2", "movbe", "f16c", "lzcnt" if "lzcnt" in flags else "abm"}
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    level = "v4" if v3 <= flags and v4 <= flags else ("v3" if v3 <= flags else None)
    return model, level


_ref_fast = None


def ref_fast():
    """The reference's code built with its own optimisation flags
    (-Ofast -funroll-loops ..., oracle/_ref/libnbpg_ref_fast_<isa>.so) for
    the ISA level this host supports: (library, ISA level), or (None, None).
    For timing only: -Ofast may reorder its float paths, so parity checks
    use ref()."""
    global _ref_fast
    if _ref_fast is None:
        _, level = host_cpu()
        path = os.path.join(HERE, "_ref", "libnbpg_ref_fast_%s.so" % level) if level else None
        if path is None or not os.path.exists(path):
            return None, None
        _ref_fast = (_bind_ref(C.CDLL(path)), level)
    return _ref_fast


def ref():
    """The reference's own code (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and ref_available():
        _ref = _bind_ref(C.CDLL(os.path.join(HERE, "_ref", "libnbpg_ref.so")))
    return _ref


def _bind_ref(R):
    """argument types of the reference shim's entry points"""
    R.nbpgr_distance.restype = C.c_float
    R.nbpgr_distance.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32]
    R.nbpgr_hash64_2.restype = C.c_uint64
    R.nbpgr_hash64_2.argtypes = [C.c_uint64]
    R.nbpgr_hash_filter_size.restype = C.c_uint32
    R.nbpgr_hash_filter_size.argtypes = [C.c_uint32]
    R.nbpgr_search.restype = C.c_int
    R.nbpgr_search.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p,
                               C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                               C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p,
                               C.c_void_p]
    R.nbpgr_search_exact.restype = C.c_int
    R.nbpgr_search_exact.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_uint32,
                                     C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                     C.c_int, C.c_void_p, C.c_uint32, C.c_void_p,
                                     C.c_uint32, C.c_uint32, C.c_double, C.c_uint64,
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                     C.c_void_p]
    R.nbpgr_repair.restype = C.c_uint64
    R.nbpgr_repair.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32]
    R.nbpgr_prune.restype = C.c_uint64
    R.nbpgr_prune.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                              C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.c_double]
    R.nbpgr_prune_batch.restype = C.c_double
    R.nbpgr_prune_batch.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p,
                                    C.c_uint64, C.c_void_p, C.c_uint32, C.c_uint32,
                                    C.c_double, C.c_double, C.c_void_p]
    R.nbpgr_shuffle.restype = None
    R.nbpgr_shuffle.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32]
    R.nbpgr_matmul.restype = None
    R.nbpgr_matmul.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                               C.c_uint32, C.c_void_p]
    R.nbpgr_compress.restype = None
    R.nbpgr_compress.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_uint32,
                                 C.c_void_p]
    R.nbpgr_open.restype = C.c_void_p
    R.nbpgr_open.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint32]
    R.nbpgr_timed_search.restype = C.c_double
    R.nbpgr_timed_search.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                     C.c_uint32, C.c_double, C.c_uint64, C.c_uint32,
                                     C.c_uint32, C.c_void_p, C.c_void_p]
    R.nbpgr_close.restype = None
    R.nbpgr_close.argtypes = [C.c_void_p]
    R.nbpgr_threads.restype = C.c_int
    R.nbpgr_threads.argtypes = []
    return R


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def points(arr, dtype):
    """Points view over a C-contiguous 2-D array (rows are packed bytes)."""
    a = np.ascontiguousarray(arr)
    esz = 4 if (dtype & 0xFF) == F32 else 1
    P = Points(a.ctypes.data, a.shape[1] * esz, a.shape[0], a.shape[1], dtype)
    P._keep = a
    return P


def graph_rows(nbr_lists, R):
    """Reference-layout rows [deg, ids...] from a list of neighbour lists."""
    n = len(nbr_lists)
    rows = np.zeros((n, R + 1), dtype=np.uint32)
    for i, l in enumerate(nbr_lists):
        rows[i, 0] = len(l)
        rows[i, 1:1 + len(l)] = l
    return rows


def distance(dtype, a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return lib().nbpgo_distance(dtype, _p(a), _p(b), a.shape[-1])


def hash64_2(x):
    return lib().nbpgo_hash64_2(int(x))


def hash_filter_size(L):
    return lib().nbpgo_hash_filter_size(int(L))


def compgt(base, queries, k, dtype=U8, tie_mode=0, nthreads=0):
    base = np.ascontiguousarray(base)
    queries = np.ascontiguousarray(queries)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32)
    d = np.zeros((nq, k), np.float32)
    rc = lib().nbpgo_compgt(dtype, _p(base), base.shape[0], _p(queries), nq, base.shape[1], k,
                            tie_mode, _p(ids), _p(d), nthreads)
    if rc != 0:
        raise ValueError("oracle compgt failed: %d" % rc)
    return ids, d


def prune(base, dtype, p, cand_ids, cand_dists, R, alpha, cos_angle=0.0):
    """sort_candidates + robustPrune; returns the kept (ids, dists)."""
    n = len(cand_ids)
    c = np.zeros(n, dtype=[("id", np.uint32), ("dist", np.float32)])
    c["id"] = cand_ids
    c["dist"] = cand_dists
    P = points(base, dtype)
    L = lib()
    m = L.nbpgo_sort_candidates(p, _p(c), n)
    mm = C.c_uint64(m)
    L.nbpgo_robust_prune(p, _p(c), C.byref(mm), C.byref(P), R, alpha, cos_angle, 1)
    m = mm.value
    return c["id"][:m].copy(), c["dist"][:m].copy()


def search(rows, base, queries, dtype, L, k, cut=1.35, limit=None, degree_limit=None,
           entry=0, max_visited=None, nthreads=0):
    """Deterministic beam search (reference semantics) for every query."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    n, R1 = rows.shape
    G = Graph(rows.ctypes.data, n, R1 - 1)
    B = points(base, dtype)
    Q = points(queries, dtype)
    nq = Q.n
    qp = QParams(k, cut, L, n if limit is None else limit,
                 (R1 - 1) if degree_limit is None else degree_limit,
                 8 * L if max_visited is None else max_visited)
    sp = np.array([entry], np.uint32)
    ids = np.zeros((nq, k), np.uint32)
    d = np.zeros((nq, k), np.float32)
    st = np.zeros((nq, 2), np.uint32)
    rc = lib().nbpgo_search(C.byref(G), C.byref(B), C.byref(Q), _p(sp), 1, C.byref(qp), _p(ids),
                            _p(d), _p(st), nthreads)
    if rc != 0:
        raise ValueError("oracle search failed: %d" % rc)
    return ids, d, st


def search_exact(rows, base, queries, dtype, raw_base, raw_queries, raw_dtype, L, k,
                 rerank_factor, cut=1.35, limit=None, degree_limit=None, entry=0, nthreads=0):
    """Beam search over quantized rows with raw rows for the final distances
    (rerank_factor > 0: rerank mode, 0: filtering mode); (ids, dists, stats)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    n, R1 = rows.shape
    G = Graph(rows.ctypes.data, n, R1 - 1)
    B, Q = points(base, dtype), points(queries, dtype)
    XB, XQ = points(raw_base, raw_dtype), points(raw_queries, raw_dtype)
    nq = Q.n
    qp = QParams(k, cut, L, n if limit is None else limit,
                 (R1 - 1) if degree_limit is None else degree_limit, 8 * L)
    sp = np.array([entry], np.uint32)
    ids = np.zeros((nq, k), np.uint32)
    d = np.zeros((nq, k), np.float32)
    st = np.zeros((nq, 2), np.uint32)
    rc = lib().nbpgo_search_exact(C.byref(G), C.byref(B), C.byref(Q), C.byref(XB), C.byref(XQ),
                                  _p(sp), 1, C.byref(qp), rerank_factor, _p(ids), _p(d), _p(st),
                                  nthreads)
    if rc != 0:
        raise ValueError("oracle search failed: %d" % rc)
    return ids, d, st


def repair(rows, sp=0):
    """DFS::repair_isolate_node restated: (repaired rows copy, isolated count)."""
    rows = np.array(rows, dtype=np.uint32, copy=True, order="C")
    G = Graph(rows.ctypes.data, rows.shape[0], rows.shape[1] - 1)
    cnt = lib().nbpgo_repair_isolated(C.byref(G), sp)
    return rows, int(cnt)


def build_graph(base, dtype, R, L, alpha=1.2, passes=1, num_split=0, VL=140, min_alpha=1.0,
                dynamic_prune=False, rand_sp=False, repair=True, top_degree=0, sp_seed=1,
                rows=None, nthreads=0, min_cos_angle=0.0, max_cos_angle=0.0):
    """build_vamana restated; returns reference-layout rows (n, R+1)."""
    base = np.ascontiguousarray(base)
    n = base.shape[0]
    if rows is None:
        rows = np.zeros((n, R + 1), np.uint32)
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    G = Graph(rows.ctypes.data, n, R)
    P = points(base, dtype)
    bp = BuildParams(R, L, VL, top_degree, num_split, passes, n, min_alpha, alpha, min_cos_angle,
                     max_cos_angle, int(dynamic_prune), int(rand_sp), int(repair), sp_seed)
    iso = C.c_uint64(0)
    rc = lib().nbpgo_build(C.byref(P), C.byref(G), C.byref(bp), C.byref(iso), nthreads)
    if rc != 0:
        raise ValueError("oracle build failed: %d" % rc)
    return rows, iso.value


def shuffle(n, seed):
    a = np.arange(n, dtype=np.uint32)
    lib().nbpgo_shuffle_mt19937(_p(a), n, seed)
    return a


def sp_stream(seed, counter):
    return lib().nbpgo_sp_stream(seed, counter)


def matmul(x, W, bias, nthreads=0):
    x = np.ascontiguousarray(x, np.float32)
    W = np.ascontiguousarray(W, np.float32)
    bias = np.ascontiguousarray(bias, np.float32)
    y = np.zeros((x.shape[0], W.shape[0]), np.float32)
    lib().nbpgo_matmul(_p(x), x.shape[0], x.shape[1], _p(W), _p(bias), W.shape[0], _p(y), nthreads)
    return y


def compress(y, is_signed, chunk, nthreads=0):
    y = np.ascontiguousarray(y, np.float32)
    out = np.zeros((y.shape[0], y.shape[1] // chunk), np.uint8)
    lib().nbpgo_compress(_p(y), y.shape[0], y.shape[1], int(is_signed), chunk, _p(out), nthreads)
    return out


def recall(result_ids, gt_ids, base, queries, dtype, k):
    result_ids = np.ascontiguousarray(result_ids, np.uint32)
    gt_ids = np.ascontiguousarray(gt_ids, np.uint32)
    B = points(base, dtype)
    Q = points(queries, dtype)
    return lib().nbpgo_recall(_p(result_ids), result_ids.shape[1], result_ids.shape[0],
                              _p(gt_ids), gt_ids.shape[1], C.byref(B), C.byref(Q), k)
