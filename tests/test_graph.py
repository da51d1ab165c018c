"""CSR graph layer: loaders, writer, generators, relabelling (CPU only)."""
import os

import numpy as np
import pytest


def check_csr(off, adj):
    n = len(off) - 1
    assert off[0] == 0 and off[-1] == len(adj)
    assert np.all(np.diff(off) >= 0)
    for v in range(n):
        row = adj[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0)  # sorted, no duplicates
        assert not np.any(row == v)  # no self-loops
    # symmetric
    src = np.repeat(np.arange(n), np.diff(off))
    a = set(zip(src.tolist(), adj.tolist()))
    assert all((w, v) in a for v, w in a)


def test_generators_are_seeded_and_valid(gat):
    for make in (lambda s: gat.Graph.erdos_renyi(500, 0.02, s),
                 lambda s: gat.Graph.rmat(10, 8, 0.57, 0.19, 0.19, s),
                 lambda s: gat.Graph.chung_lu(2000, 10.0, 2.1, 100, s)):
        g1, g2, g3 = make(1), make(1), make(2)
        o1, a1 = g1.csr()
        o2, a2 = g2.csr()
        o3, a3 = g3.csr()
        check_csr(o1, a1)
        assert np.array_equal(o1, o2) and np.array_equal(a1, a2)
        assert not (len(a1) == len(a3) and np.array_equal(a1, a3))


def test_er_expected_edge_count(gat):
    # C1: G(n=10000, p=16/9999): E[edges] = n(n-1)/2 * p = 80000, sd ~ 283
    g = gat.Graph.erdos_renyi(10000, 16 / 9999, 1)
    assert abs(g.undirected_edges() - 80000) < 5 * 283
    assert g.v_cnt() == 10000


def test_rmat_and_chung_lu_shape(gat):
    g = gat.Graph.rmat(12, 16, 0.57, 0.19, 0.19, 20)
    assert g.v_cnt() == 4096
    assert g.undirected_edges() <= 16 * 4096
    deg = np.diff(g.csr()[0])
    assert deg[0] > np.median(deg) * 10  # skew toward low ids (a = 0.57)
    cl = gat.Graph.chung_lu(20000, 24.0, 2.1, 500, 21)
    d = np.diff(cl.csr()[0])
    assert d.max() <= 2 * 500  # expected-degree cap, Poisson tail
    assert 15 < d.mean() < 26


def test_bin_roundtrip_and_layout(gat, tmp_path):
    g = gat.Graph.rmat(8, 4, 0.57, 0.19, 0.19, 5)
    p = tmp_path / "g.bin"
    g.save(p)
    raw = p.read_bytes()
    n = np.frombuffer(raw[:4], dtype=np.int32)[0]
    m = np.frombuffer(raw[4:12], dtype=np.int64)[0]
    off = np.frombuffer(raw[12:12 + 8 * (n + 1)], dtype=np.int64)
    adj = np.frombuffer(raw[12 + 8 * (n + 1):], dtype=np.int32)
    o, a = g.csr()
    assert n == g.v_cnt() and m == g.e_cnt() and np.array_equal(off, o) and np.array_equal(adj, a)
    h = gat.Graph.load(p, binary=True)
    o2, a2 = h.csr()
    assert np.array_equal(o, o2) and np.array_equal(a, a2)


def test_bin_rejects_truncated(gat, tmp_path):
    from graphautotune_amd._lib import GatError
    g = gat.Graph.erdos_renyi(50, 0.1, 1)
    p = tmp_path / "g.bin"
    g.save(p)
    p.write_bytes(p.read_bytes()[:-4])
    with pytest.raises(GatError):
        gat.Graph.load(p)


def test_bin_rejects_asymmetric(gat, tmp_path):
    # a sorted, loop-free .bin whose entry (0,2) has no reverse (2,0): the
    # engine's orientation and shards assume symmetric rows
    from graphautotune_amd._lib import GatError
    off = np.array([0, 2, 3, 4], dtype=np.int64)
    adj = np.array([1, 2, 0, 1], dtype=np.int32)
    p = tmp_path / "asym.bin"
    p.write_bytes(np.int32(3).tobytes() + np.int64(4).tobytes() + off.tobytes() + adj.tobytes())
    with pytest.raises(GatError, match="symmetric"):
        gat.Graph.load(p)
    # rows {1,2}, {0}, {0}: symmetric
    adj2 = np.array([1, 2, 0, 0], dtype=np.int32)
    p.write_bytes(np.int32(3).tobytes() + np.int64(4).tobytes() + off.tobytes() + adj2.tobytes())
    assert gat.Graph.load(p).e_cnt() == 4


def test_txt_loader_semantics(gat, tmp_path):
    # graph_backend.cuh:23-89: relabel by sorted id, drop self-loops,
    # symmetrise, dedup; declared count may exceed the ids seen.
    p = tmp_path / "g.txt"
    p.write_text("6 6\n10 20\n20 10\n30 30\n20 40\n10 40\n40 20\n")
    g = gat.Graph.load(p, binary=False)
    o, a = g.csr()
    assert g.v_cnt() == 6
    # ids 10,20,30,40 -> 0,1,2,3; edges {0-1, 1-3, 0-3}
    rows = [a[o[v]:o[v + 1]].tolist() for v in range(6)]
    assert rows == [[1, 3], [0, 3], [], [0, 1], [], []]
    q = tmp_path / "bad.txt"
    q.write_text("2 2\n1 2\n3 4\n")
    from graphautotune_amd._lib import GatError
    with pytest.raises(GatError):
        gat.Graph.load(q, binary=False)


def test_from_csr_validation(gat):
    from graphautotune_amd._lib import GatError
    ok = gat.Graph.from_csr([0, 1, 2], [1, 0])
    assert ok.e_cnt() == 2
    with pytest.raises(GatError):
        gat.Graph.from_csr([0, 1, 1], [1])  # not symmetric
    with pytest.raises(GatError):
        gat.Graph.from_csr([0, 2, 2, 2], [2, 1])  # unsorted row


def test_degree_histogram(gat):
    g = gat.Graph.erdos_renyi(300, 0.05, 2)
    h = g.degree_histogram()
    deg = np.diff(g.csr()[0])
    want = np.bincount(np.floor(np.log2(deg + 1)).astype(int), minlength=32)[:32]
    assert np.array_equal(h, want) and h.sum() == 300


def test_triangle_count(gat, orc):
    from patterns import TRIANGLE
    for g in (gat.Graph.erdos_renyi(300, 0.1, 5), gat.Graph.rmat(9, 8, 0.57, 0.19, 0.19, 4),
              gat.Graph.chung_lu(400, 10.0, 2.1, 60, 2)):
        off, adj = g.csr()
        assert g.triangles() == orc.unique_count(off, adj, TRIANGLE)
    m = 23  # K_m: C(m,3)
    src = [i for i in range(m) for j in range(i + 1, m)]
    dst = [j for i in range(m) for j in range(i + 1, m)]
    assert gat.Graph.from_edges(m, src, dst).triangles() == m * (m - 1) * (m - 2) // 6
    assert gat.Graph.from_edges(4, [0, 1, 2], [1, 2, 3]).triangles() == 0


def test_golden_clique_graphs_unchanged(gat):
    """The graphs behind tests/golden/clique_counts.json are still the ones
    the generator makes (fingerprints of their CSR arrays)."""
    import json
    import os
    import zlib
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "clique_counts.json")) as f:
        gold = json.load(f)
    for scale in (12, 16, 20):
        g = gat.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, 20)  # csr() views g's memory: keep g alive
        off, adj = g.csr()
        want = gold["rmat%d" % scale]
        assert zlib.adler32(np.ascontiguousarray(off).tobytes()) == want["adler32_offsets"]
        assert zlib.adler32(np.ascontiguousarray(adj).tobytes()) == want["adler32_adj"]
