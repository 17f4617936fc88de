"""Host-side mirror of the reference's search interface over the C ABI.

    Vectors        VectoresGPU + subirVectores / copiarVectoresAsync   (cuda/datos_gpu.h:8-14)
    FlatGraph      GrafoGPU + subirGrafo                               (cuda/datos_gpu.h:19-28)
    HnswIndex      HnswGPU + leerHnsw/subirHnsw                        (cuda/hnsw.h:43-62)
    SearchConfig   ConfigBusqueda                                      (cuda/busqueda.cuh:6-11)
    Searcher       BuscadorGPU: prepare / launch / free                (cuda/busqueda.cuh:36-62)
    HnswSearcher   BuscadorHnsw: prepare / launch / free               (cuda/hnsw.h:65-86)

Everything runs through libpann.so; there is no other compute path.  Errors raise
PannError (the reference exits the process instead).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import PannError, check, lib

NO_NEIGHBOR = _lib.NO_NEIGHBOR
FLAG_OVERFLOW = _lib.FLAG_OVERFLOW
FLAG_SHORT = _lib.FLAG_SHORT
FLAG_CONFIG = _lib.FLAG_CONFIG


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def device_count():
    n = C.c_int(0)
    code = lib().pann_device_count(C.byref(n))
    return n.value if code == 0 else 0


def synchronize(stream=None):
    check(lib().pann_synchronize(stream))


class DeviceBuffer:
    """Plain device allocation (pann_device_alloc)."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(lib().pann_device_alloc(nbytes, C.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    @staticmethod
    def from_array(a, stream=None):
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(a.nbytes)
        check(lib().pann_memcpy(b.ptr, _ptr(a), a.nbytes, stream))
        return b

    def to_array(self, dtype, shape, stream=None):
        out = np.empty(shape, dtype=dtype)
        check(lib().pann_memcpy(_ptr(out), self.ptr, out.nbytes, stream))
        return out

    def free(self):
        if self.ptr:
            lib().pann_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Vectors:
    """Device rows with a 128-byte pitch (VectoresGPU)."""

    def __init__(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int8)
        h = C.c_void_p()
        check(lib().pann_vectors_create(_ptr(rows), rows.shape[0], rows.shape[1], C.byref(h)))
        self.handle = h.value
        self.n, self.dim = rows.shape

    @classmethod
    def load(cls, path):
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(lib().pann_vectors_load(str(path).encode(), C.byref(h)))
        self.handle = h.value
        n, d = C.c_uint32(), C.c_uint32()
        check(lib().pann_vectors_info(self.handle, C.byref(n), C.byref(d), None, None))
        self.n, self.dim = n.value, d.value
        return self

    def device_rows(self):
        p = C.c_void_p()
        check(lib().pann_vectors_info(self.handle, None, None, C.byref(p), None))
        return p.value

    def copy_async(self, rows, stream=None):
        rows = np.ascontiguousarray(rows, dtype=np.int8)
        check(lib().pann_vectors_copy_async(self.handle, _ptr(rows), rows.shape[0], stream))

    def free(self):
        if self.handle:
            lib().pann_vectors_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FlatGraph:
    """Dense n x max_degree adjacency in HBM (GrafoGPU)."""

    def __init__(self, graph):
        h = C.c_void_p()
        deg = np.ascontiguousarray(graph.degrees, dtype=np.uint32)
        edges = np.ascontiguousarray(graph.edges, dtype=np.uint32)
        check(lib().pann_graph_create(graph.n, graph.max_degree, _ptr(deg), _ptr(edges), C.byref(h)))
        self.handle = h.value
        self.n, self.max_degree = graph.n, max(1, graph.max_degree)

    @classmethod
    def load(cls, path):
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(lib().pann_graph_load(str(path).encode(), C.byref(h)))
        self.handle = h.value
        n, g = C.c_uint32(), C.c_uint32()
        check(lib().pann_graph_info(self.handle, C.byref(n), C.byref(g)))
        self.n, self.max_degree = n.value, g.value
        return self

    def free(self):
        if self.handle:
            lib().pann_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HnswIndex:
    """HNSW layers in HBM with vectors in node order (HnswGPU)."""

    def __init__(self, h, base):
        out = C.c_void_p()
        arrs = [np.ascontiguousarray(h.level, dtype=np.uint32),
                np.ascontiguousarray(h.node_point, dtype=np.uint32),
                np.ascontiguousarray(h.first_list, dtype=np.uint64),
                np.ascontiguousarray(h.list_offsets, dtype=np.uint64),
                np.ascontiguousarray(h.list_sizes, dtype=np.uint32),
                np.ascontiguousarray(h.list_ids, dtype=np.uint32),
                np.ascontiguousarray(h.entries, dtype=np.uint32)]
        check(lib().pann_hnsw_create(h.n, *[_ptr(a) for a in arrs[:6]], _ptr(arrs[6]),
                                     len(h.entries), base.handle, C.byref(out)))
        self._keep = arrs
        self.handle = out.value
        self.n = h.n

    @classmethod
    def load(cls, path, base):
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(lib().pann_hnsw_load(str(path).encode(), base.handle, C.byref(h)))
        self.handle = h.value
        n = C.c_uint32()
        check(lib().pann_hnsw_info(self.handle, C.byref(n), None, None, None, None))
        self.n = n.value
        return self

    def free(self):
        if self.handle:
            lib().pann_hnsw_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class SearchConfig:
    """ConfigBusqueda: Ls, visit limit (None = n), k, cut."""
    Ls: int = 10
    k: int = 10
    visit_limit: int = None
    cut: float = 1.35

    def c(self):
        return _lib.SearchConfig(self.Ls, self.k, int(self.visit_limit or 0), float(self.cut))


@dataclass
class Results:
    ids: np.ndarray
    dists: np.ndarray
    comps: np.ndarray
    flags: np.ndarray
    edges: np.ndarray
    hops: np.ndarray


class _SearcherBase:
    handle = None

    def outputs(self):
        o = _lib.SearchOutputs()
        check(lib().pann_searcher_outputs(self.handle, C.byref(o)))
        return o

    def download(self, stream=None):
        """Copies the last launch's outputs to host arrays (blocking when stream is None)."""
        nq, k = self.nq, max(1, self.cfg.k)
        r = Results(np.empty((nq, k), np.uint32), np.empty((nq, k), np.int32),
                    np.empty(nq, np.int32), np.empty(nq, np.int32), np.empty(nq, np.int32),
                    np.empty(nq, np.int32))
        check(lib().pann_searcher_download(self.handle, _ptr(r.ids), _ptr(r.dists), _ptr(r.comps),
                                           _ptr(r.flags), _ptr(r.edges), _ptr(r.hops), stream))
        if stream is not None:
            synchronize(stream)
        r.ids, r.dists = r.ids[:, :self.cfg.k], r.dists[:, :self.cfg.k]
        return r

    def free(self):
        if self.handle:
            lib().pann_searcher_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Searcher(_SearcherBase):
    """BuscadorGPU: prepare(graph, nq, cfg) once, launch(...) many times."""

    def __init__(self, graph, nq, cfg):
        h = C.c_void_p()
        cc = cfg.c()
        check(lib().pann_searcher_prepare(graph.handle, nq, C.byref(cc), C.byref(h)))
        self.handle = h.value
        self.nq, self.cfg = nq, cfg

    def launch(self, base, queries, graph, qids=None, stream=None):
        """Asynchronous on `stream`; qids is an optional DeviceBuffer of nq u32."""
        check(lib().pann_searcher_launch(self.handle, base.handle, queries.handle, graph.handle,
                                         qids.ptr if qids is not None else None, stream))


class HnswSearcher(_SearcherBase):
    """BuscadorHnsw: prepare(index, nq, cfg) once, launch(...) many times."""

    def __init__(self, index, nq, cfg):
        h = C.c_void_p()
        cc = cfg.c()
        check(lib().pann_hnsw_searcher_prepare(index.handle, nq, C.byref(cc), C.byref(h)))
        self.handle = h.value
        self.nq, self.cfg = nq, cfg

    def launch(self, queries, index, stream=None):
        check(lib().pann_hnsw_searcher_launch(self.handle, queries.handle, index.handle, stream))


def search_flat(base, graph, queries, cfg, qids=None):
    """One blocking search: returns Results with ids/dists of shape [nq, k]."""
    s = Searcher(graph, queries.n, cfg)
    try:
        qb = DeviceBuffer.from_array(np.asarray(qids, dtype=np.uint32)) if qids is not None else None
        s.launch(base, queries, graph, qb)
        synchronize()
        return s.download()
    finally:
        s.free()


PARAMS_DTYPE = np.dtype([("qid", "<u4"), ("Ls", "<i4"), ("k", "<i4"), ("reserved", "<i4"),
                         ("visit_limit", "<i8"), ("cut", "<f8")])


def make_params(nq, Ls, k, visit_limit=None, cut=1.35, qids=None):
    """pann_query_params records: scalars broadcast, arrays per query."""
    p = np.zeros(nq, dtype=PARAMS_DTYPE)
    p["qid"] = np.arange(nq) if qids is None else qids
    p["Ls"] = Ls
    p["k"] = k
    p["visit_limit"] = 0 if visit_limit is None else visit_limit
    p["cut"] = cut
    return p


class Batch:
    """Batched persistent form: queries with their own qid/Ls/k/visit/cut drained by one
    launch of resident wavefronts; optionally captured into a HIP graph and replayed."""

    def __init__(self, max_queries, max_Ls, max_k, base=None, graph=None, hnsw=None,
                 waves_per_block=0, lds_filter_bits=0, visited_slots=0, prefetch=0, config=None,
                 waves_per_query=0):
        cfg = config or _lib.BatchConfig(max_Ls, max_k, waves_per_block, lds_filter_bits,
                                          visited_slots, prefetch, waves_per_query)
        h = C.c_void_p()
        if hnsw is not None:
            check(lib().pann_batch_prepare_hnsw(hnsw.handle, max_queries, C.byref(cfg), C.byref(h)))
        else:
            check(lib().pann_batch_prepare_flat(base.handle, graph.handle, max_queries,
                                                C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self.max_k = cfg.max_k
        self.nq = 0
        self._keep = None

    def set_queries(self, rows, params, stream=None):
        rows = np.ascontiguousarray(rows, dtype=np.int8)
        params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
        self._keep = (rows, params)  # alive until the asynchronous copies complete
        check(lib().pann_batch_set_queries(self.handle, _ptr(rows), _ptr(params), len(params), stream))
        self.nq = len(params)

    def launch(self, stream=None):
        check(lib().pann_batch_launch(self.handle, stream))

    def capture(self, stream):
        check(lib().pann_batch_capture(self.handle, stream))

    def replay(self, stream=None):
        check(lib().pann_batch_replay(self.handle, stream))

    def config(self):
        c, w = _lib.BatchConfig(), C.c_uint32()
        check(lib().pann_batch_info(self.handle, C.byref(c), C.byref(w)))
        return {f: getattr(c, f) for f, _ in c._fields_}, w.value

    def download(self, stream=None):
        nq, k = self.nq, self.max_k
        r = Results(np.empty((nq, k), np.uint32), np.empty((nq, k), np.int32),
                    np.empty(nq, np.int32), np.empty(nq, np.int32), np.empty(nq, np.int32),
                    np.empty(nq, np.int32))
        check(lib().pann_batch_download(self.handle, _ptr(r.ids), _ptr(r.dists), _ptr(r.comps),
                                        _ptr(r.flags), _ptr(r.edges), _ptr(r.hops), stream))
        if stream is not None:
            synchronize(stream)
        return r

    def free(self):
        if self.handle:
            lib().pann_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def autotune(rows, params, max_Ls, max_k, base=None, graph=None, hnsw=None):
    """Fastest clean batch configuration on this sample: (BatchConfig, ms per launch)."""
    rows = np.ascontiguousarray(rows, dtype=np.int8)
    params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
    best, ms = _lib.BatchConfig(), C.c_float()
    check(lib().pann_batch_autotune(base.handle if base is not None else None,
                                    graph.handle if graph is not None else None,
                                    hnsw.handle if hnsw is not None else None, _ptr(rows),
                                    _ptr(params), len(params), max_Ls, max_k, C.byref(best),
                                    C.byref(ms)))
    return best, ms.value


@dataclass
class Recorded:
    """Served queries recorded by the load generator, in service order."""
    ids: np.ndarray     # [m, 10]: first min(k, 10) ids, NO_NEIGHBOR after
    dists: np.ndarray   # [m, 10]
    flags: np.ndarray   # [m]
    rows: np.ndarray    # [m]: row of the query set (also the query's qid)
    params: np.ndarray  # [m] PARAMS_DTYPE


class Server:
    """Continuous-batching serving loop over a resident index (pann_server_*)."""

    def __init__(self, slots, max_batch, streams, max_Ls, max_k, base=None, graph=None, hnsw=None,
                 waves_per_block=0, lds_filter_bits=0, visited_slots=0):
        cfg = _lib.ServerConfig(slots, max_batch, streams, max_Ls, max_k, waves_per_block,
                                lds_filter_bits, visited_slots)
        h = C.c_void_p()
        check(lib().pann_server_create(base.handle if base is not None else None,
                                       graph.handle if graph is not None else None,
                                       hnsw.handle if hnsw is not None else None, C.byref(cfg),
                                       C.byref(h)))
        self.handle = h.value
        self.max_k = max_k

    def submit(self, rows, params):
        rows = np.ascontiguousarray(rows, dtype=np.int8)
        params = np.ascontiguousarray(params, dtype=PARAMS_DTYPE)
        rid = C.c_uint64()
        check(lib().pann_server_submit(self.handle, _ptr(rows), _ptr(params), len(params), C.byref(rid)))
        return rid.value, len(params)

    def wait(self, ticket):
        rid, nq = ticket
        ids = np.empty((nq, self.max_k), np.uint32)
        d = np.empty((nq, self.max_k), np.int32)
        f = np.empty(nq, np.int32)
        ms = C.c_double()
        check(lib().pann_server_wait(self.handle, rid, _ptr(ids), _ptr(d), _ptr(f), C.byref(ms)))
        return ids, d, f, ms.value

    def stats(self):
        s = _lib.ServerStats()
        check(lib().pann_server_stats_get(self.handle, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def run_poisson(self, rows, rate_qps, duration_s, max_request=4096, Ls_lo=10, Ls_hi=300,
                    ks=(10, 100), seed=5, record_cap=0, full_record=False):
        """Open-loop Poisson load.  Returns (stats dict, recorded ids [m, 10], recorded rows),
        or with full_record=True (stats, Recorded) holding ids, dists [m, 10], flags, rows
        and the parameters of the first m <= record_cap served queries."""
        rows = np.ascontiguousarray(rows, dtype=np.int8)
        ks = np.ascontiguousarray(ks, dtype=np.int32)
        s = _lib.ServerStats()
        cap = max(1, record_cap)
        rec = Recorded(np.empty((cap, 10), np.uint32), np.empty((cap, 10), np.int32),
                       np.empty(cap, np.int32), np.empty(cap, np.uint32),
                       np.zeros(cap, PARAMS_DTYPE))
        got = C.c_uint32()
        on = (lambda a: _ptr(a) if record_cap else None)
        check(lib().pann_server_run_poisson(self.handle, _ptr(rows), len(rows), float(rate_qps),
                                            float(duration_s), max_request, Ls_lo, Ls_hi, _ptr(ks),
                                            len(ks), seed, C.byref(s), on(rec.ids), on(rec.dists),
                                            on(rec.flags), on(rec.rows), on(rec.params),
                                            record_cap, C.byref(got)))
        st = {f: getattr(s, f) for f, _ in s._fields_}
        m = got.value
        rec = Recorded(rec.ids[:m], rec.dists[:m], rec.flags[:m], rec.rows[:m], rec.params[:m])
        if full_record:
            return st, rec
        return st, rec.ids, rec.rows

    def free(self):
        if self.handle:
            lib().pann_server_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().pann_stream_create(C.byref(s)))
        self.ptr = s.value

    def synchronize(self):
        synchronize(self.ptr)

    def free(self):
        if self.ptr:
            lib().pann_stream_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gather_bench(rows, queries, count, seed=1, reps=3):
    """Random-gather ceiling over the rows of `rows` (pann_gather_bench): returns (mean ms
    per launch, checksum, gathered GB/s counting 128 bytes per distance)."""
    ms, cs = C.c_float(), C.c_uint64()
    check(lib().pann_gather_bench(rows.handle, queries.handle, int(count), int(seed), int(reps),
                                  C.byref(ms), C.byref(cs)))
    return ms.value, cs.value, 128.0 * count / (ms.value * 1e-3) / 1e9


def bruteforce_knn(base, queries, k, nq=None, stream=None):
    """Exact top-k on the GPU (MFMA distances, fused selection): ids u32, dists i32 [nq, k]."""
    nq = queries.n if nq is None else nq
    ids = DeviceBuffer(max(1, nq * k * 4))
    d = DeviceBuffer(max(1, nq * k * 4))
    try:
        check(lib().pann_bruteforce_knn(base.handle, queries.handle, nq, k, ids.ptr, d.ptr, stream))
        return ids.to_array(np.uint32, (nq, k)), d.to_array(np.int32, (nq, k))
    finally:
        ids.free(), d.free()


__all__ = ["bruteforce_knn", "Vectors", "FlatGraph", "HnswIndex", "SearchConfig", "Searcher", "HnswSearcher",
           "DeviceBuffer", "Results", "search_flat", "device_count", "synchronize", "PannError",
           "NO_NEIGHBOR", "FLAG_OVERFLOW", "FLAG_SHORT", "FLAG_CONFIG"]
