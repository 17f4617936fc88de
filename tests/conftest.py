import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "parlayann-gpu-busqueda_amd")
for p in (PKG, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def has_gpu():
    from pann import search
    try:
        return search.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.skip("no HIP device")
    return True


# ---- the 10k x 100 case: 32-component mixture, Vamana R=32 L=64 alpha=1.2 ----
@pytest.fixture(scope="session")
def w1():
    from pann import builders, datagen
    base, queries = datagen.gmm_dataset(10000, 1000, dim=100, ncomp=32)
    graph = builders.build_vamana(base, R=32, L=64, alpha=1.2, two_pass=True, seed=7)
    return base, queries, graph


@pytest.fixture(scope="session")
def small():
    """2,000 x 100 points, 200 queries, Vamana R=24 L=48."""
    from pann import builders, datagen
    base, queries = datagen.gmm_dataset(2000, 200, dim=100, ncomp=8, seeds=(11, 12, 13))
    graph = builders.build_vamana(base, R=24, L=48, alpha=1.2, two_pass=True, seed=3)
    return base, queries, graph


def random_graph(n, max_degree, rng, dup_frac=0.0, min_degree=1):
    """Random directed graph; dup_frac of the edges repeat an earlier edge of the row."""
    from pann import formats
    deg = rng.integers(min_degree, max_degree + 1, size=n).astype(np.uint32)
    rows = []
    for i in range(n):
        r = rng.integers(0, n, size=int(deg[i])).astype(np.uint32)
        if dup_frac > 0 and deg[i] > 1:
            m = rng.random(int(deg[i])) < dup_frac
            m[0] = False
            src = rng.integers(0, np.arange(int(deg[i])).clip(min=1))
            r[m] = r[src[m]]
        rows.append(r)
    return formats.Graph(n, max_degree, deg, np.concatenate(rows))
