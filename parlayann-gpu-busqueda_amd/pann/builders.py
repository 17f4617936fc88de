"""Graph builders (libpann_tools.so) returning the on-disk structures of pann.formats."""
import ctypes as C
import os
import tempfile

import numpy as np

from . import formats
from ._lib import check_tools, tools


def _nthreads(n=None):
    if n:
        return int(n)
    return int(os.environ.get("OMP_NUM_THREADS") or min(16, os.cpu_count() or 1))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def build_vamana(base, R=32, L=64, alpha=1.2, two_pass=True, seed=7, nthreads=None):
    base = np.ascontiguousarray(base, dtype=np.int8)
    n, dim = base.shape
    deg = np.zeros(n, dtype=np.uint32)
    dense = np.zeros((n, R), dtype=np.uint32)
    check_tools(tools().pann_build_vamana(_ptr(base), n, dim, R, L, float(alpha), int(two_pass),
                                          seed, _nthreads(nthreads), _ptr(deg), _ptr(dense)))
    return formats.Graph.from_dense(deg, dense, R)


def build_vamana_gpu(vectors, R=32, L=64, alpha=1.2, two_pass=True, seed=7):
    """build_vamana's graph, built on the device from the rows of a search.Vectors
    (pann_build_vamana_gpu); equal to build_vamana's for the same arguments."""
    from ._lib import check, lib
    n = vectors.n
    deg = np.zeros(n, dtype=np.uint32)
    dense = np.zeros((n, R), dtype=np.uint32)
    check(lib().pann_build_vamana_gpu(vectors.handle, R, L, float(alpha), int(two_pass), seed,
                                      _ptr(deg), _ptr(dense), None))
    return formats.Graph.from_dense(deg, dense, R)


def _via_file(fn, suffix):
    fd, path = tempfile.mkstemp(suffix=suffix)
    os.close(fd)
    try:
        fn(path.encode())
        return path
    except BaseException:
        os.unlink(path)
        raise


def build_hnsw(base, M=32, efc=128, ml=None, alpha=1.0, seed=7, nthreads=None, path=None):
    base = np.ascontiguousarray(base, dtype=np.int8)
    n, dim = base.shape
    ml = float(ml if ml is not None else 1.0 / np.log(M))

    def run(p):
        check_tools(tools().pann_build_hnsw(_ptr(base), n, dim, M, efc, ml, float(alpha), seed,
                                            _nthreads(nthreads), p))
    if path:
        run(str(path).encode())
        return formats.read_hnsw(path)
    tmp = _via_file(run, ".hnsw")
    try:
        return formats.read_hnsw(tmp)
    finally:
        os.unlink(tmp)


def build_hnsw_gpu(vectors, M=32, efc=128, ml=None, alpha=1.0, seed=7, path=None):
    """build_hnsw's index, built on the device from the rows of a search.Vectors
    (pann_build_hnsw_gpu); its v3 file equals build_hnsw's for the same arguments."""
    from ._lib import check, lib
    ml = float(ml if ml is not None else 1.0 / np.log(M))

    def run(p):
        check(lib().pann_build_hnsw_gpu(vectors.handle, M, efc, ml, float(alpha), seed, p, None))
    if path:
        run(str(path).encode())
        return formats.read_hnsw(path)
    tmp = _via_file(run, ".hnsw")
    try:
        return formats.read_hnsw(tmp)
    finally:
        os.unlink(tmp)


def build_hcnng(base, num_trees=30, leaf_size=1000, mst_deg=3, seed=7, nthreads=None, path=None):
    base = np.ascontiguousarray(base, dtype=np.int8)
    n, dim = base.shape

    def run(p):
        check_tools(tools().pann_build_hcnng(_ptr(base), n, dim, num_trees, leaf_size, mst_deg,
                                             seed, _nthreads(nthreads), p))
    if path:
        run(str(path).encode())
        return formats.read_graph(path)
    tmp = _via_file(run, ".graph")
    try:
        return formats.read_graph(tmp)
    finally:
        os.unlink(tmp)


def build_hcnng_gpu(vectors, num_trees=30, leaf_size=1000, mst_deg=3, seed=7, path=None):
    """HCNNG built on the device over the rows of a search.Vectors (pann_build_hcnng_gpu);
    its graph file equals build_hcnng's for the same arguments."""
    from ._lib import check, lib

    def run(p):
        check(lib().pann_build_hcnng_gpu(vectors.handle, num_trees, leaf_size, mst_deg, seed, p, None))
    if path:
        run(str(path).encode())
        return formats.read_graph(path)
    tmp = _via_file(run, ".graph")
    try:
        return formats.read_graph(tmp)
    finally:
        os.unlink(tmp)


def build_pynndescent(base, K=40, num_trees=10, leaf_size=1000, alpha=1.2, delta=0.05,
                      max_degree=None, seed=7, nthreads=None, path=None):
    base = np.ascontiguousarray(base, dtype=np.int8)
    n, dim = base.shape
    md = int(max_degree or K)

    def run(p):
        check_tools(tools().pann_build_pynndescent(_ptr(base), n, dim, K, num_trees, leaf_size,
                                                   float(alpha), float(delta), md, seed,
                                                   _nthreads(nthreads), p))
    if path:
        run(str(path).encode())
        return formats.read_graph(path)
    tmp = _via_file(run, ".graph")
    try:
        return formats.read_graph(tmp)
    finally:
        os.unlink(tmp)


def build_pynndescent_gpu(vectors, K=40, num_trees=10, leaf_size=1000, alpha=1.2, delta=0.05,
                          max_degree=None, seed=7, path=None):
    """PyNNDescent built on the device over the rows of a search.Vectors
    (pann_build_pynndescent_gpu); its graph file equals build_pynndescent's for the same
    arguments."""
    from ._lib import check, lib
    md = int(max_degree or K)

    def run(p):
        check(lib().pann_build_pynndescent_gpu(vectors.handle, K, num_trees, leaf_size, float(alpha),
                                               float(delta), md, seed, p, None))
    if path:
        run(str(path).encode())
        return formats.read_graph(path)
    tmp = _via_file(run, ".graph")
    try:
        return formats.read_graph(tmp)
    finally:
        os.unlink(tmp)
