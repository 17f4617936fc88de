"""ctypes binding of the C ABI in include/pann.h (libpann.so) and include/pann_tools.h
(libpann_tools.so).  The libraries are built in-tree by build.py; loading fails loudly
when they are missing."""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PANN_PROFILE=1 selects the section-profiling build (build.py --profile)
LIB_PATH = os.path.join(PKG_ROOT, "libpann_prof.so" if os.environ.get("PANN_PROFILE") == "1"
                        else "libpann.so")
TOOLS_PATH = os.path.join(PKG_ROOT, "libpann_tools.so")

PANN_OK = 0
PANN_ERR_ARG = 1
PANN_ERR_IO = 2
PANN_ERR_FORMAT = 3
PANN_ERR_HIP = 4
PANN_ERR_CAPACITY = 5
NO_NEIGHBOR = 0xFFFFFFFF
FLAG_OVERFLOW = 1
FLAG_SHORT = 2
FLAG_CONFIG = 4


class PannError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pann error {code}: {msg}")
        self.code = code


class SearchConfig(C.Structure):
    """ConfigBusqueda (cuda/busqueda.cuh:6-11)."""
    _fields_ = [("Ls", C.c_int32), ("k", C.c_int32), ("visit_limit", C.c_int64),
                ("cut", C.c_double)]


class QueryParams(C.Structure):
    _fields_ = [("qid", C.c_uint32), ("Ls", C.c_int32), ("k", C.c_int32),
                ("reserved", C.c_int32), ("visit_limit", C.c_int64), ("cut", C.c_double)]


class SearchOutputs(C.Structure):
    _fields_ = [("ids", C.c_void_p), ("dists", C.c_void_p), ("comps", C.c_void_p),
                ("flags", C.c_void_p), ("edges", C.c_void_p), ("hops", C.c_void_p),
                ("stride", C.c_uint32), ("nq", C.c_uint32)]


class BatchConfig(C.Structure):
    _fields_ = [("max_Ls", C.c_int32), ("max_k", C.c_int32), ("waves_per_block", C.c_int32),
                ("lds_filter_bits", C.c_int32), ("visited_slots", C.c_uint32),
                ("prefetch", C.c_uint32), ("waves_per_query", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("max_batch", C.c_uint32), ("streams", C.c_int32),
                ("max_Ls", C.c_int32), ("max_k", C.c_int32), ("waves_per_block", C.c_int32),
                ("lds_filter_bits", C.c_int32), ("visited_slots", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("queries", C.c_uint64), ("launches", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("max_ms", C.c_double), ("busy_s", C.c_double)]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_u32 = C.c_uint32
_i32 = C.c_int32
_u64 = C.c_uint64

# name -> (restype, argtypes); every symbol declared in include/pann.h
PANN_SIGNATURES = {
    "pann_last_error": (C.c_char_p, []),
    "pann_version": (C.c_int, []),
    "pann_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pann_synchronize": (C.c_int, [_P]),
    "pann_device_synchronize": (C.c_int, []),
    "pann_device_alloc": (C.c_int, [C.c_size_t, _PP]),
    "pann_device_free": (None, [_P]),
    "pann_host_alloc": (C.c_int, [C.c_size_t, _PP]),
    "pann_host_free": (None, [_P]),
    "pann_memcpy": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "pann_stream_create": (C.c_int, [_PP]),
    "pann_stream_destroy": (None, [_P]),
    "pann_event_create": (C.c_int, [_PP]),
    "pann_event_record": (C.c_int, [_P, _P]),
    "pann_event_elapsed_ms": (C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    "pann_event_destroy": (None, [_P]),
    "pann_copy_kernel": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "pann_gather_bench": (C.c_int, [_P, _P, _u64, _u64, C.c_int, C.POINTER(C.c_float),
                                    C.POINTER(_u64)]),
    "pann_profile_counters": (C.c_int, [_P, C.c_int, C.c_int]),
    "pann_vectors_create": (C.c_int, [_P, _u32, _u32, _PP]),
    "pann_vectors_load": (C.c_int, [C.c_char_p, _PP]),
    "pann_vectors_copy_async": (C.c_int, [_P, _P, _u32, _P]),
    "pann_vectors_info": (C.c_int, [_P, C.POINTER(_u32), C.POINTER(_u32), _PP, C.POINTER(_u32)]),
    "pann_vectors_free": (None, [_P]),
    "pann_graph_create": (C.c_int, [_u32, _u32, _P, _P, _PP]),
    "pann_graph_load": (C.c_int, [C.c_char_p, _PP]),
    "pann_graph_info": (C.c_int, [_P, C.POINTER(_u32), C.POINTER(_u32)]),
    "pann_graph_free": (None, [_P]),
    "pann_hnsw_load": (C.c_int, [C.c_char_p, _P, _PP]),
    "pann_hnsw_create": (C.c_int, [_u32, _P, _P, _P, _P, _P, _P, _P, _u32, _P, _PP]),
    "pann_hnsw_info": (C.c_int, [_P, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32),
                                 C.POINTER(_u32), C.POINTER(_u32)]),
    "pann_hnsw_free": (None, [_P]),
    "pann_searcher_prepare": (C.c_int, [_P, _u32, C.POINTER(SearchConfig), _PP]),
    "pann_searcher_launch": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "pann_search_flat_noalloc": (C.c_int, [_P, _P, _P, _P, _u32, C.POINTER(SearchConfig), _P, _P,
                                           _P, _P, _P, _P, _P]),
    "pann_hnsw_searcher_prepare": (C.c_int, [_P, _u32, C.POINTER(SearchConfig), _PP]),
    "pann_hnsw_searcher_launch": (C.c_int, [_P, _P, _P, _P]),
    "pann_searcher_outputs": (C.c_int, [_P, C.POINTER(SearchOutputs)]),
    "pann_searcher_download": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "pann_searcher_free": (None, [_P]),
    "pann_bruteforce_knn": (C.c_int, [_P, _P, _u32, _u32, _P, _P, _P]),
    "pann_build_vamana_gpu": (C.c_int, [_P, _u32, _u32, C.c_float, C.c_int, _u64, _P, _P, _P]),
    "pann_build_hnsw_gpu": (C.c_int, [_P, _u32, _u32, C.c_float, C.c_float, _u64, C.c_char_p, _P]),
    "pann_build_hcnng_gpu": (C.c_int, [_P, _u32, _u32, _u32, _u64, C.c_char_p, _P]),
    "pann_build_pynndescent_gpu": (C.c_int, [_P, _u32, _u32, _u32, C.c_float, C.c_float, _u32, _u64,
                                             C.c_char_p, _P]),
    "pann_batch_prepare_flat": (C.c_int, [_P, _P, _u32, C.POINTER(BatchConfig), _PP]),
    "pann_batch_prepare_hnsw": (C.c_int, [_P, _u32, C.POINTER(BatchConfig), _PP]),
    "pann_batch_set_queries": (C.c_int, [_P, _P, _P, _u32, _P]),
    "pann_batch_launch": (C.c_int, [_P, _P]),
    "pann_batch_capture": (C.c_int, [_P, _P]),
    "pann_batch_replay": (C.c_int, [_P, _P]),
    "pann_batch_outputs": (C.c_int, [_P, C.POINTER(SearchOutputs)]),
    "pann_batch_download": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "pann_batch_info": (C.c_int, [_P, C.POINTER(BatchConfig), C.POINTER(_u32)]),
    "pann_batch_free": (None, [_P]),
    "pann_batch_autotune": (C.c_int, [_P, _P, _P, _P, _P, _u32, _i32, _i32,
                                      C.POINTER(BatchConfig), C.POINTER(C.c_float)]),
    "pann_server_create": (C.c_int, [_P, _P, _P, C.POINTER(ServerConfig), _PP]),
    "pann_server_submit": (C.c_int, [_P, _P, _P, _u32, C.POINTER(_u64)]),
    "pann_server_wait": (C.c_int, [_P, _u64, _P, _P, _P, C.POINTER(C.c_double)]),
    "pann_server_stats_get": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "pann_server_free": (None, [_P]),
    "pann_server_run_poisson": (C.c_int, [_P, _P, _u32, C.c_double, C.c_double, _u32, _i32, _i32,
                                          _P, _u32, _u64, C.POINTER(ServerStats), _P, _P, _P,
                                          _P, _P, _u32, C.POINTER(_u32)]),
}

TOOLS_SIGNATURES = {
    "pann_tools_last_error": (C.c_char_p, []),
    "pann_build_vamana": (C.c_int, [_P, _u32, _u32, _u32, _u32, C.c_float, C.c_int, _u64,
                                    C.c_int, _P, _P]),
    "pann_build_hnsw": (C.c_int, [_P, _u32, _u32, _u32, _u32, C.c_float, C.c_float, _u64,
                                  C.c_int, C.c_char_p]),
    "pann_build_hcnng": (C.c_int, [_P, _u32, _u32, _u32, _u32, _u32, _u64, C.c_int,
                                   C.c_char_p]),
    "pann_build_pynndescent": (C.c_int, [_P, _u32, _u32, _u32, _u32, _u32, C.c_float,
                                         C.c_float, _u32, _u64, C.c_int, C.c_char_p]),
    "pann_write_graph": (C.c_int, [C.c_char_p, _u32, _u32, _P, _P, _u32]),
}

_lib = None
_tools = None


def _bind(path, sigs):
    if not os.path.exists(path):
        raise PannError(PANN_ERR_IO, f"{path} is not built (run build.py or __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(lib, name, None)
        if f is None:
            continue  # calling it raises AttributeError; tests check the export list
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    global _lib
    if _lib is None:
        _lib = _bind(LIB_PATH, PANN_SIGNATURES)
    return _lib


def tools():
    global _tools
    if _tools is None:
        _tools = _bind(TOOLS_PATH, TOOLS_SIGNATURES)
    return _tools


def check(code):
    if code != PANN_OK:
        raise PannError(code, lib().pann_last_error().decode(errors="replace"))


def check_tools(code):
    if code != PANN_OK:
        raise PannError(code, tools().pann_tools_last_error().decode(errors="replace"))
