"""File formats: round trips through the numpy writers and both readers (numpy and the
library's C++ readers behind pann_*_load), the reference's size/range checks, and the
neighbour file the reference itself wrote."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import random_graph

HERE = os.path.dirname(os.path.abspath(__file__))


def test_i8bin_roundtrip(tmp_path):
    from pann import formats
    x = np.random.default_rng(0).integers(-128, 128, size=(17, 100)).astype(np.int8)
    p = tmp_path / "x.i8bin"
    formats.write_i8bin(p, x)
    assert os.path.getsize(p) == 8 + 17 * 100
    assert np.array_equal(formats.read_i8bin(p), x)
    with open(p, "ab") as f:
        f.write(b"\0")
    with pytest.raises(ValueError):
        formats.read_i8bin(p)


def test_graph_roundtrip_and_checks(tmp_path):
    from pann import formats
    g = random_graph(50, 9, np.random.default_rng(1), dup_frac=0.2)
    p = tmp_path / "g"
    formats.write_graph(p, g)
    h = formats.read_graph(p)
    assert h.n == 50 and h.max_degree == 9
    assert np.array_equal(h.degrees, g.degrees) and np.array_equal(h.edges, g.edges)
    d = h.dense()
    assert d.shape == (50, 9)
    assert np.array_equal(formats.Graph.from_dense(h.degrees, d).edges, g.edges)
    # truncated
    raw = open(p, "rb").read()
    open(tmp_path / "t", "wb").write(raw[:-4])
    with pytest.raises(ValueError):
        formats.read_graph(tmp_path / "t")
    # degree above max degree
    bad = bytearray(raw)
    bad[4:8] = (2).to_bytes(4, "little")
    open(tmp_path / "m", "wb").write(bad)
    with pytest.raises(ValueError, match="max degree"):
        formats.read_graph(tmp_path / "m")
    # edge out of range
    bad = bytearray(raw)
    bad[-4:] = (50).to_bytes(4, "little")
    open(tmp_path / "e", "wb").write(bad)
    with pytest.raises(ValueError, match="points to"):
        formats.read_graph(tmp_path / "e")


def test_gt_and_neighbors_roundtrip(tmp_path):
    from pann import formats
    rng = np.random.default_rng(2)
    ids = rng.integers(0, 1000, size=(7, 100)).astype(np.uint32)
    d = rng.random((7, 100)).astype(np.float32)
    formats.write_gt(tmp_path / "gt", ids, d)
    i2, d2 = formats.read_gt(tmp_path / "gt")
    assert np.array_equal(i2, ids) and np.array_equal(d2, d)
    formats.write_neighbors(tmp_path / "nb.bin", ids[:, :10])
    assert np.array_equal(formats.read_neighbors(tmp_path / "nb.bin"), ids[:, :10])


def test_reference_neighbor_file_fixture():
    """ParDiskANN_Ls_75.bin as written by the reference (cuda/comparacion1.cu:126-138):
    29,316 queries x 10, ids of the 1M MSSPACEV base, 130 missing neighbours."""
    from pann import formats
    a = formats.read_neighbors(os.path.join(HERE, "golden", "ParDiskANN_Ls_75.bin"))
    assert a.shape == (29316, 10)
    miss = a == 0xFFFFFFFF
    assert int(miss.sum()) == 130
    assert int(a[~miss].max()) < 1_000_000
    # missing neighbours only ever trail a row (short frontiers)
    rows = np.flatnonzero(miss.any(1))
    for r in rows:
        first = int(np.argmax(miss[r]))
        assert miss[r, first:].all()


def _hnsw_example():
    from pann import formats
    lv = np.array([0, 2, 1, 0, 0], np.uint32)
    pts = np.array([4, 3, 2, 1, 0], np.uint32)
    lists = {(0, 0): [1, 2], (1, 0): [0, 2, 3], (1, 1): [2], (1, 2): [], (2, 0): [1, 4],
             (2, 1): [1], (3, 0): [1], (4, 0): [2, 3]}
    first, offs, sizes, ids = [], [], [], []
    for i in range(5):
        first.append(len(offs))
        for layer in range(int(lv[i]) + 1):
            offs.append(len(ids))
            sizes.append(len(lists[(i, layer)]))
            ids += lists[(i, layer)]
    return formats.Hnsw(5, 100, 32, 128, 0.36, 0.83, lv, pts, np.array(first, np.uint64),
                        np.array(offs, np.uint64), np.array(sizes, np.uint32),
                        np.array(ids, np.uint32), np.array([1], np.uint32))


def test_hnsw_roundtrip_and_checks(tmp_path):
    from pann import formats
    h = _hnsw_example()
    p = tmp_path / "h"
    formats.write_hnsw(p, h)
    r = formats.read_hnsw(p)
    for layer_node in [(0, 0), (1, 0), (1, 1), (1, 2), (2, 1), (4, 0)]:
        assert list(r.neighbors(layer_node[0], layer_node[1])) == list(h.neighbors(*layer_node))
    assert list(r.entries) == [1] and list(r.node_point) == [4, 3, 2, 1, 0]
    raw = open(p, "rb").read()
    open(tmp_path / "tr", "wb").write(raw + b"\0")
    with pytest.raises(ValueError, match="trailing"):
        formats.read_hnsw(tmp_path / "tr")
    bad = bytearray(raw)
    bad[4:8] = (2).to_bytes(4, "little")
    open(tmp_path / "v", "wb").write(bad)
    with pytest.raises(ValueError, match="version"):
        formats.read_hnsw(tmp_path / "v")


def test_library_readers_report_errors_without_a_device(tmp_path):
    """pann_graph_load / pann_hnsw_load / pann_vectors_load validate files on the host
    and return PANN_ERR_IO / PANN_ERR_FORMAT (the reference exits instead)."""
    from pann import _lib, formats
    L = _lib.lib()
    out = C.c_void_p()
    assert L.pann_graph_load(str(tmp_path / "missing").encode(), C.byref(out)) == _lib.PANN_ERR_IO
    g = random_graph(20, 4, np.random.default_rng(3))
    formats.write_graph(tmp_path / "g", g)
    raw = open(tmp_path / "g", "rb").read()
    open(tmp_path / "g2", "wb").write(raw[:-1])
    assert L.pann_graph_load(str(tmp_path / "g2").encode(), C.byref(out)) == _lib.PANN_ERR_FORMAT
    assert b"size" in L.pann_last_error()
    formats.write_hnsw(tmp_path / "h", _hnsw_example())
    open(tmp_path / "h2", "wb").write(open(tmp_path / "h", "rb").read() + b"\0\0")
    assert L.pann_hnsw_load(str(tmp_path / "h2").encode(), None, C.byref(out)) == _lib.PANN_ERR_FORMAT
    assert b"trailing" in L.pann_last_error()
    open(tmp_path / "x.i8bin", "wb").write(b"\x02\0\0\0\x03\0\0\0" + b"\0" * 5)
    assert L.pann_vectors_load(str(tmp_path / "x.i8bin").encode(), C.byref(out)) == _lib.PANN_ERR_FORMAT
