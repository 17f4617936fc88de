"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline leg
of bench.py, as the checker.  Function-by-function citations and what pins each one
are in oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class QueryParams(C.Structure):
    _fields_ = [("qid", C.c_uint32), ("Ls", C.c_int32), ("k", C.c_int32), ("pad_", C.c_int32),
                ("visit_limit", C.c_int64), ("cut", C.c_double)]


class HnswC(C.Structure):
    _fields_ = [("n", C.c_uint32), ("level", C.c_void_p), ("node_point", C.c_void_p),
                ("node_first_list", C.c_void_p), ("list_offsets", C.c_void_p),
                ("list_sizes", C.c_void_p), ("list_ids", C.c_void_p), ("entries", C.c_void_p),
                ("num_entries", C.c_uint32)]


PARAMS_DTYPE = np.dtype([("qid", "<u4"), ("Ls", "<i4"), ("k", "<i4"), ("pad_", "<i4"),
                         ("visit_limit", "<i8"), ("cut", "<f8")])

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.oracle_hash64_2.restype = C.c_uint64
        L.oracle_hash64_2.argtypes = [C.c_uint64]
        L.oracle_hash64.restype = C.c_uint64
        L.oracle_hash64.argtypes = [C.c_uint64]
        L.oracle_dist_i8.restype = C.c_int32
        L.oracle_dist_i8.argtypes = [P, P, C.c_uint32]
        L.oracle_beam_flat_batch.restype = C.c_int
        L.oracle_beam_flat_batch.argtypes = [P, C.c_uint32, C.c_uint32, P, P, P, P, C.c_uint32,
                                             P, C.c_int, C.c_uint32, P, P, P, P, P, P, C.c_int]
        L.oracle_search_hnsw_batch.restype = C.c_int
        L.oracle_search_hnsw_batch.argtypes = [P, C.c_uint32, C.POINTER(HnswC), P, C.c_uint32,
                                               P, C.c_int, C.c_uint32, P, P, P, P, P, P, C.c_int]
        L.oracle_bruteforce_knn.restype = C.c_int
        L.oracle_bruteforce_knn.argtypes = [P, C.c_uint32, C.c_uint32, P, C.c_uint32,
                                            C.c_uint32, P, P, C.c_int]
        L.oracle_recall.restype = C.c_double
        L.oracle_recall.argtypes = [P, P, C.c_uint32, C.c_uint32, P, C.c_uint32, C.c_uint32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _threads(t):
    return int(t or min(16, os.cpu_count() or 1))


def hash64_2(x):
    return lib().oracle_hash64_2(int(x))


def hash64(x):
    return lib().oracle_hash64(int(x))


def dist(a, b):
    a = np.ascontiguousarray(a, dtype=np.int8)
    b = np.ascontiguousarray(b, dtype=np.int8)
    return lib().oracle_dist_i8(_p(a), _p(b), a.shape[0])


def make_params(nq, Ls, k, visit_limit=None, cut=1.35, qids=None):
    p = np.zeros(nq, dtype=PARAMS_DTYPE)
    p["qid"] = np.arange(nq) if qids is None else qids
    p["Ls"] = Ls
    p["k"] = k
    p["visit_limit"] = 0 if visit_limit is None else visit_limit
    p["cut"] = cut
    return p


class Result:
    def __init__(self, ids, dists, comps, flags, edges, hops):
        self.ids, self.dists, self.comps, self.flags, self.edges, self.hops = (
            ids, dists, comps, flags, edges, hops)


def _outputs(nq, stride):
    return (np.zeros((nq, stride), np.uint32), np.zeros((nq, stride), np.int32),
            np.zeros(nq, np.int32), np.zeros(nq, np.int32), np.zeros(nq, np.int64),
            np.zeros(nq, np.int32))


def beam_flat(base, graph, queries, params, nthreads=None):
    """ParlayANN flat beam search for every query; params: PARAMS_DTYPE array [nq]."""
    base = np.ascontiguousarray(base, dtype=np.int8)
    queries = np.ascontiguousarray(queries, dtype=np.int8)
    params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
    nq = queries.shape[0]
    stride = max(1, int(params["k"].max()) if nq else 1)
    deg = np.ascontiguousarray(graph.degrees, dtype=np.uint32)
    off = np.ascontiguousarray(graph.offsets, dtype=np.uint64)
    edges = np.ascontiguousarray(graph.edges, dtype=np.uint32)
    if edges.size == 0:
        edges = np.zeros(1, np.uint32)
    ids, dists, comps, flags, ed, hops = _outputs(nq, stride)
    r = lib().oracle_beam_flat_batch(_p(base), base.shape[0], base.shape[1], _p(deg), _p(off),
                                     _p(edges), _p(queries), nq, _p(params), 0, stride, _p(ids),
                                     _p(dists), _p(comps), _p(flags), _p(ed), _p(hops),
                                     _threads(nthreads))
    if r:
        raise ValueError("oracle_beam_flat_batch rejected its arguments")
    return Result(ids, dists, comps, flags, ed, hops)


def search_hnsw(base, h, queries, params, nthreads=None):
    base = np.ascontiguousarray(base, dtype=np.int8)
    queries = np.ascontiguousarray(queries, dtype=np.int8)
    params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
    nq = queries.shape[0]
    stride = max(1, int(params["k"].max()) if nq else 1)
    keep = [np.ascontiguousarray(h.level, np.uint32), np.ascontiguousarray(h.node_point, np.uint32),
            np.ascontiguousarray(h.first_list, np.uint64),
            np.ascontiguousarray(h.list_offsets, np.uint64),
            np.ascontiguousarray(h.list_sizes, np.uint32),
            np.ascontiguousarray(h.list_ids if h.list_ids.size else np.zeros(1), np.uint32),
            np.ascontiguousarray(h.entries, np.uint32)]
    hc = HnswC(h.n, *[a.ctypes.data for a in keep], len(h.entries))
    ids, dists, comps, flags, ed, hops = _outputs(nq, stride)
    r = lib().oracle_search_hnsw_batch(_p(base), base.shape[1], C.byref(hc), _p(queries), nq,
                                       _p(params), 0, stride, _p(ids), _p(dists), _p(comps),
                                       _p(flags), _p(ed), _p(hops), _threads(nthreads))
    if r:
        raise ValueError("oracle_search_hnsw_batch rejected its arguments")
    return Result(ids, dists, comps, flags, ed, hops)


def bruteforce_knn(base, queries, k, nthreads=None):
    base = np.ascontiguousarray(base, dtype=np.int8)
    queries = np.ascontiguousarray(queries, dtype=np.int8)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32)
    d = np.zeros((nq, k), np.int32)
    r = lib().oracle_bruteforce_knn(_p(base), base.shape[0], base.shape[1], _p(queries), nq, k,
                                    _p(ids), _p(d), _threads(nthreads))
    if r:
        raise ValueError("k larger than n")
    return ids, d


def recall(gt_ids, gt_dists, ids, k):
    gt_ids = np.ascontiguousarray(gt_ids, dtype=np.uint32)
    gt_dists = np.ascontiguousarray(gt_dists, dtype=np.float32)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    return lib().oracle_recall(_p(gt_ids), _p(gt_dists), gt_ids.shape[0], gt_ids.shape[1],
                               _p(ids), ids.shape[1], k)
