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
