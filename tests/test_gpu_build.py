"""GPU Vamana construction (pann_build_vamana_gpu) against the host builder
(pann_build_vamana, the batch insertion of ParlayANN's Vamana): the same degrees and the
same rows, entry for entry, for the same arguments.  Cases cover one point, frontier
widths of one, two and four registers (L 16 / 100 / 200), rows wider than 64 (R = 100),
rows with and without their norm embedded (dim 100 vs 128), single and two-pass builds,
and duplicated points (distance-0 ties broken by id)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # n, dim, R, L, alpha, two_pass, seed
    (1, 16, 8, 8, 1.2, True, 1),
    (300, 16, 8, 16, 1.2, True, 3),
    (3000, 100, 32, 64, 1.2, True, 7),
    (3000, 128, 24, 100, 1.3, False, 5),
    (6000, 37, 64, 200, 1.2, True, 11),
    (2500, 32, 100, 128, 1.1, True, 2),
]


def _data(n, dim, seed):
    from pann import datagen
    base, _ = datagen.gmm_dataset(n, 1, dim=dim, ncomp=max(1, n // 200), seeds=(seed, seed + 1, seed + 2))
    if n >= 100:
        base[n // 2: n // 2 + 10] = base[:10]   # duplicated points
    return base


@pytest.mark.parametrize("n,dim,R,L,alpha,two_pass,seed", CASES)
def test_gpu_build_equals_host(gpu, n, dim, R, L, alpha, two_pass, seed):
    from pann import builders, search
    base = _data(n, dim, seed)
    print(f"[build] n={n} dim={dim} R={R} L={L} alpha={alpha} two_pass={two_pass}", flush=True)
    host = builders.build_vamana(base, R=R, L=L, alpha=alpha, two_pass=two_pass, seed=seed, nthreads=8)
    b = search.Vectors(base)
    try:
        g = builders.build_vamana_gpu(b, R=R, L=L, alpha=alpha, two_pass=two_pass, seed=seed)
    finally:
        b.free()
    assert np.array_equal(g.degrees, host.degrees)
    assert np.array_equal(g.edges, host.edges)


def test_gpu_build_rejects_bad_arguments(gpu):
    from pann import PannError, builders, search
    b = search.Vectors(np.zeros((4, 8), np.int8))
    try:
        with pytest.raises(PannError):
            builders.build_vamana_gpu(b, R=129, L=16)
        with pytest.raises(PannError):
            builders.build_vamana_gpu(b, R=8, L=257)
    finally:
        b.free()


HNSW_CASES = [
    # n, dim, M, efc, alpha, seed
    (1, 16, 4, 8, 1.0, 1),
    (400, 16, 4, 16, 1.0, 3),
    (4000, 100, 16, 64, 1.0, 7),
    (4000, 128, 8, 100, 1.2, 5),
    (6000, 37, 32, 200, 1.0, 11),
]


@pytest.mark.parametrize("n,dim,M,efc,alpha,seed", HNSW_CASES)
def test_gpu_hnsw_build_equals_host(gpu, tmp_path, n, dim, M, efc, alpha, seed):
    """The HNSW v3 file written by the device build equals the host builder's, byte for
    byte (levels, every layer's lists, entry point); ml = 1/ln M."""
    from pann import builders, search
    base = _data(n, dim, seed)
    print(f"[build] hnsw n={n} dim={dim} M={M} efc={efc} alpha={alpha}", flush=True)
    builders.build_hnsw(base, M=M, efc=efc, alpha=alpha, seed=seed, nthreads=8, path=tmp_path / "host.hnsw")
    b = search.Vectors(base)
    try:
        builders.build_hnsw_gpu(b, M=M, efc=efc, alpha=alpha, seed=seed, path=tmp_path / "gpu.hnsw")
    finally:
        b.free()
    assert (tmp_path / "gpu.hnsw").read_bytes() == (tmp_path / "host.hnsw").read_bytes()


HCNNG_CASES = [
    # n, dim, num_trees, leaf_size, mst_deg, seed
    (2, 16, 1, 2, 1, 1),
    (500, 16, 3, 40, 3, 3),
    (4000, 100, 4, 300, 3, 7),
    (5000, 128, 3, 1000, 2, 5),
    (6000, 37, 5, 1024, 4, 11),
    (3000, 100, 2, 17, 3, 2),
]


@pytest.mark.parametrize("n,dim,trees,leaf,deg,seed", HCNNG_CASES)
def test_gpu_hcnng_build_equals_host(gpu, tmp_path, n, dim, trees, leaf, deg, seed):
    """The HCNNG graph file written by the device build equals the host builder's, byte
    for byte: the same cluster-tree leaves, the same 16 nearest leaf mates, the same
    degree-bounded spanning forests and the same order of every node's list.  Cases cover
    leaves of 2 points, leaves of 17 to 1,024 points, rows with and without their norm
    embedded (dim 100 vs 128), degree bounds 1 to 4 and duplicated points."""
    from pann import builders, search
    base = _data(n, dim, seed)
    print(f"[build] hcnng n={n} dim={dim} trees={trees} leaf={leaf} deg={deg}", flush=True)
    builders.build_hcnng(base, num_trees=trees, leaf_size=leaf, mst_deg=deg, seed=seed, nthreads=8,
                         path=tmp_path / "host.graph")
    b = search.Vectors(base)
    try:
        builders.build_hcnng_gpu(b, num_trees=trees, leaf_size=leaf, mst_deg=deg, seed=seed,
                                 path=tmp_path / "gpu.graph")
    finally:
        b.free()
    assert (tmp_path / "gpu.graph").read_bytes() == (tmp_path / "host.graph").read_bytes()


def test_gpu_hcnng_rejects_bad_arguments(gpu):
    from pann import PannError, builders, search
    b = search.Vectors(np.zeros((4, 8), np.int8))
    try:
        with pytest.raises(PannError):
            builders.build_hcnng_gpu(b, num_trees=1, leaf_size=1025, mst_deg=3)
        with pytest.raises(PannError):
            builders.build_hcnng_gpu(b, num_trees=1, leaf_size=100, mst_deg=5)
    finally:
        b.free()


PYNN_CASES = [
    # n, dim, K, num_trees, leaf_size, alpha, delta, max_degree, seed, ncomp
    (2, 16, 4, 1, 2, 1.2, 0.05, 4, 1, 1),
    (600, 16, 8, 3, 50, 1.2, 0.05, 8, 3, 3),
    (4000, 100, 16, 4, 300, 1.2, 0.05, 16, 7, 20),
    (5000, 128, 40, 3, 1000, 1.2, 0.05, 40, 5, 25),
    (6000, 37, 24, 5, 1024, 1.1, 0.02, 32, 11, 30),
    (3000, 100, 40, 2, 200, 1.3, 0.1, 60, 2, 15),
]


@pytest.mark.parametrize("n,dim,K,trees,leaf,alpha,delta,md,seed,ncomp", PYNN_CASES)
def test_gpu_pynndescent_build_equals_host(gpu, tmp_path, n, dim, K, trees, leaf, alpha, delta, md, seed,
                                           ncomp):
    """The PyNNDescent graph file written by the device build equals the host builder's,
    byte for byte: the same cluster-tree K-NN lists, the same NN-descent rounds (sampled
    new/old lists, reverse lists, joins, stop rule), the same robust prune of lists plus
    reverse edges, and the same leaf bridges between components.  Cases cover K from 4 to
    40, lists shorter than K (leaves of 2 points), rows with and without their norm
    embedded, clustered data with many components, and duplicated points."""
    from pann import builders, datagen, search
    base, _ = datagen.gmm_dataset(n, 1, dim=dim, ncomp=ncomp, seeds=(seed, seed + 1, seed + 2))
    if n >= 100:
        base[n // 2: n // 2 + 10] = base[:10]
    print(f"[build] pynndescent n={n} dim={dim} K={K} trees={trees} leaf={leaf} md={md}", flush=True)
    builders.build_pynndescent(base, K=K, num_trees=trees, leaf_size=leaf, alpha=alpha, delta=delta,
                               max_degree=md, seed=seed, nthreads=8, path=tmp_path / "host.graph")
    b = search.Vectors(base)
    try:
        builders.build_pynndescent_gpu(b, K=K, num_trees=trees, leaf_size=leaf, alpha=alpha, delta=delta,
                                       max_degree=md, seed=seed, path=tmp_path / "gpu.graph")
    finally:
        b.free()
    assert (tmp_path / "gpu.graph").read_bytes() == (tmp_path / "host.graph").read_bytes()


def test_gpu_pynndescent_rejects_bad_arguments(gpu):
    from pann import PannError, builders, search
    b = search.Vectors(np.zeros((4, 8), np.int8))
    try:
        with pytest.raises(PannError):
            builders.build_pynndescent_gpu(b, K=65)
        with pytest.raises(PannError):
            builders.build_pynndescent_gpu(b, K=8, leaf_size=2000)
        with pytest.raises(PannError):
            builders.build_pynndescent_gpu(b, K=8, max_degree=129)
    finally:
        b.free()
