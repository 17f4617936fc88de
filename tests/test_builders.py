"""Host graph builders: structural invariants, determinism across thread counts, and that
the graphs they write are valid inputs of the search (read back by the formats module)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def data():
    from pann import datagen
    return datagen.gmm_dataset(3000, 200, dim=100, ncomp=8, seeds=(61, 62, 63))


def _reachable(g):
    off = g.offsets
    seen = np.zeros(g.n, bool)
    seen[0] = True
    st = [0]
    while st:
        u = st.pop()
        for a in g.edges[int(off[u]):int(off[u + 1])]:
            if not seen[a]:
                seen[a] = True
                st.append(int(a))
    return seen.all()


def test_vamana_structure_and_determinism(data):
    from pann import builders
    base, _ = data
    g1 = builders.build_vamana(base, R=24, L=48, alpha=1.2, seed=3, nthreads=1)
    g4 = builders.build_vamana(base, R=24, L=48, alpha=1.2, seed=3, nthreads=4)
    assert np.array_equal(g1.degrees, g4.degrees) and np.array_equal(g1.edges, g4.edges)
    assert g1.degrees.max() <= 24 and g1.edges.max() < len(base)
    off = g1.offsets
    for i in range(0, len(base), 97):
        row = g1.edges[int(off[i]):int(off[i + 1])]
        assert i not in row and len(set(row.tolist())) == len(row)


def test_vamana_recall(w1):
    # the 10k x 100, 32-component recipe with R=32, L=64, alpha=1.2
    from oracle import oracle as O
    base, q, g = w1
    gt, gd = O.bruteforce_knn(base, q, 10)
    r = O.beam_flat(base, g, q, O.make_params(len(q), 40, 10))
    assert O.recall(gt, gd.astype(np.float32), r.ids, 10) > 0.97


def test_hnsw_structure_and_determinism(data, tmp_path):
    from pann import builders
    base, _ = data
    h1 = builders.build_hnsw(base, M=8, efc=32, seed=4, nthreads=1)
    h4 = builders.build_hnsw(base, M=8, efc=32, seed=4, nthreads=4, path=tmp_path / "h")
    assert np.array_equal(h1.list_ids, h4.list_ids) and np.array_equal(h1.level, h4.level)
    top = h1.level[h1.entries[0]]
    assert top == h1.level.max()
    for i in range(0, h1.n, 53):
        for layer in range(int(h1.level[i]) + 1):
            nb = h1.neighbors(i, layer)
            assert len(nb) <= (16 if layer == 0 else 8)
            assert (h1.level[nb] >= layer).all()


def test_hcnng_structure(data):
    from pann import builders
    base, _ = data
    g = builders.build_hcnng(base, num_trees=6, leaf_size=300, mst_deg=3, seed=5, nthreads=4)
    g1 = builders.build_hcnng(base, num_trees=6, leaf_size=300, mst_deg=3, seed=5, nthreads=1)
    assert np.array_equal(g.edges, g1.edges)
    # each tree contributes a forest of degree <= 3; bridges add at most a few edges
    assert g.degrees.max() <= 6 * 3 + 4
    assert _reachable(g)


def test_pynndescent_structure(data):
    from pann import builders
    base, _ = data
    g = builders.build_pynndescent(base, K=16, num_trees=4, leaf_size=300, max_degree=16, seed=6,
                                   nthreads=4)
    g1 = builders.build_pynndescent(base, K=16, num_trees=4, leaf_size=300, max_degree=16, seed=6,
                                    nthreads=1)
    assert np.array_equal(g.degrees, g1.degrees) and np.array_equal(g.edges, g1.edges)
    assert _reachable(g)


def test_pynndescent_clustered_data_is_searchable():
    """On a mixture whose clusters are far apart every K-NN list stays inside its cluster;
    the leaf bridges must make the graph strongly connected, so a search from node 0
    reaches every cluster (recall measured with the oracle against exact neighbours)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    from oracle import oracle as O
    from pann import builders, datagen
    base, queries = datagen.gmm_dataset(6000, 200, dim=100, ncomp=24, seeds=(31, 32, 33))
    g = builders.build_pynndescent(base, K=16, num_trees=4, leaf_size=400, seed=2, nthreads=4)
    rows = np.repeat(np.arange(g.n), g.degrees.astype(np.int64))
    m = csr_matrix((np.ones(len(g.edges)), (rows, g.edges.astype(np.int64))), shape=(g.n, g.n))
    assert connected_components(m, directed=True, connection="strong")[0] == 1
    gt_ids, gt_d = O.bruteforce_knn(base, queries, 10, nthreads=4)
    o = O.beam_flat(base, g, queries, O.make_params(len(queries), 50, 10), nthreads=4)
    assert O.recall(gt_ids, gt_d, o.ids, 10) >= 0.9
