"""GPU parity: counts from the HIP engine (through the C ABI) against the CPU
oracle on the same seeded graphs, bit for bit.

unique   == oracle labeled count / |Aut|   (backtracking, schedule-independent)
reference== oracle restatement of the reference engine for the same schedule
labeled  == oracle labeled count            (= naive_pm.cu's quantity)
"""
import itertools
import sys

import numpy as np
import pytest

from patterns import C5, HOUSE, K4, K5, P3, TRIANGLE, clique, connected_patterns

pytestmark = pytest.mark.gpu

CONFIGS = [
    {},  # engine defaults
    {"slot_capacity": 8, "queue_chunk": 1},  # tiny LDS buffers: most tasks take the HBM overflow launch
    {"queue_chunk": 64, "intersect_strategy": 1, "block_waves": 2, "waves_per_cu": 4},
    {"intersect_strategy": 2, "search_linear_threshold": 16, "gallop_ratio": 2},
    {"intersect_strategy": 3, "vertex_store": 1},  # merge path (cliques on the array store too)
    {"intersect_strategy": 4, "vertex_store": 1, "slot_capacity": 8, "queue_chunk": 1},  # galloping, HBM buffers
]


def graphs(gat):
    out = {
        "er200": gat.Graph.erdos_renyi(200, 0.08, 7),
        "rmat8": gat.Graph.rmat(8, 6, 0.57, 0.19, 0.19, 3),
        "cl300": gat.Graph.chung_lu(300, 6.0, 2.1, 40, 5),
    }
    return out


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_graphs(gat, gpu):
    return graphs(gat)


@pytest.mark.parametrize("pattern", [TRIANGLE, K4, K5, HOUSE, C5, P3])
@pytest.mark.parametrize("gname", ["er200", "rmat8", "cl300"])
def test_unique_matches_oracle(gat, orc, small_graphs, pattern, gname):
    g = small_graphs[gname]
    off, adj = g.csr()
    want = orc.unique_count(off, adj, pattern)
    plan = gat.Plan(g, [pattern], "unique")
    for cfg in CONFIGS:
        say("unique", gname, pattern, cfg)
        got, st = plan.count(cfg)
        assert got[0] == want, (gname, pattern, cfg, got[0], want)


@pytest.mark.parametrize("pattern", [TRIANGLE, K4, K5, HOUSE, C5])
@pytest.mark.parametrize("gname", ["er200", "rmat8", "cl300"])
def test_reference_matches_oracle(gat, orc, small_graphs, pattern, gname):
    from graphautotune_amd._lib import ScheduleData
    g = small_graphs[gname]
    off, adj = g.csr()
    sched = orc.schedule(pattern, ScheduleData)
    want, ok = orc.reference_count(off, adj, sched)
    assert ok
    plan = gat.Plan(g, [pattern], "reference")
    for cfg in CONFIGS:
        say("reference", gname, pattern, cfg)
        got, _ = plan.count(cfg)
        assert got[0] == want, (gname, pattern, cfg, got[0], want)


def test_labeled_matches_naive(gat, orc, gpu):
    g = gat.Graph.erdos_renyi(40, 0.25, 11)
    off, adj = g.csr()
    for p in (TRIANGLE, K4, HOUSE):
        say("labeled", p)
        want = orc.naive_count(off, adj, p)  # target/naive_pm.cu semantics
        got, _ = gat.Plan(g, [p], "labeled").count()
        assert got[0] == want


def test_reference_is_k_times_unique_for_cliques(gat, gpu):
    g = gat.Graph.rmat(10, 8, 0.57, 0.19, 0.19, 9)
    for k in (3, 4, 5):
        say("clique", k)
        u, _ = gat.Plan(g, [clique(k)], "unique").count()
        r, _ = gat.Plan(g, [clique(k)], "reference").count()
        assert r[0] == k * u[0]


def test_all_connected_5_vertex_patterns_unique(gat, orc, gpu):
    g = gat.Graph.erdos_renyi(120, 0.12, 21)
    off, adj = g.csr()
    pats = connected_patterns(5)
    want = [orc.unique_count(off, adj, p) for p in pats]
    say("batched 21x5v")
    got, _ = gat.Plan(g, pats, "unique").count()
    assert got == want
    for p, w in zip(pats, want):  # plain single-pattern plans agree too
        got1, _ = gat.Plan(g, [p], "unique").count({"slot_capacity": 16})
        assert got1[0] == w, p


def test_random_6_vertex_patterns_unique(gat, orc, gpu):
    rng = np.random.default_rng(6)
    pats6 = connected_patterns(6)
    pick = [pats6[i] for i in rng.choice(len(pats6), 12, replace=False)]
    g = gat.Graph.erdos_renyi(60, 0.2, 5)
    off, adj = g.csr()
    want = [orc.unique_count(off, adj, p) for p in pick]
    say("batched 12x6v", pick)
    got, _ = gat.Plan(g, pick, "unique").count({"slot_capacity": 32})
    assert got == want


def test_golden_intersection_15(gat, gpu):
    # tests/test_vertex_set.cu:44-48,76: |{1..30} ∩ {2,4,..,60}| = 15. As a
    # graph: a ~ 1..30, b ~ even 2..60, edge a-b; triangles = common
    # neighbours of a and b.
    a, b = 61, 62
    src = [a] * 30 + [b] * 30 + [a]
    dst = list(range(1, 31)) + [2 * i for i in range(1, 31)] + [b]
    g = gat.Graph.from_edges(63, src, dst)
    got, _ = gat.Plan(g, [TRIANGLE], "unique").count()
    assert got[0] == 15


def test_edge_cases(gat, gpu):
    empty = gat.Graph.from_edges(5, [], [])
    assert gat.Plan(empty, [TRIANGLE, K4], "unique").count()[0] == [0, 0]
    star = gat.Graph.from_edges(10, [0] * 9, list(range(1, 10)))
    assert gat.Plan(star, [TRIANGLE], "unique").count()[0] == [0]
    kk = gat.Graph.from_edges(9, *zip(*[(i, j) for i in range(9) for j in range(i + 1, 9)]))
    from math import comb
    got, _ = gat.Plan(kk, [TRIANGLE, K4, K5, C5, HOUSE], "unique").count()
    assert got == [comb(9, 3), comb(9, 4), comb(9, 5), 12 * comb(9, 5), 60 * comb(9, 5)]
    # complete bipartite K_{4,5}: no odd cycles
    bip = gat.Graph.from_edges(9, *zip(*[(i, j) for i in range(4) for j in range(4, 9)]))
    got, _ = gat.Plan(bip, [TRIANGLE, K4, C5, HOUSE], "unique").count()
    assert got == [0, 0, 0, 0]


def test_graph_replay_matches_eager(gat, gpu):
    g = gat.Graph.rmat(9, 8, 0.57, 0.19, 0.19, 3)
    plan = gat.Plan(g, [K4, HOUSE], "unique")
    eager, _ = plan.count()
    for _ in range(3):
        got, _ = plan.count(graph_replays=3)
        assert got == eager
        again, _ = plan.count()  # eager work between replays
        assert again == eager


def test_walg_matches_oracle(gat, orc, gpu):
    g = gat.Graph.rmat(9, 8, 0.57, 0.19, 0.19, 3)
    off, adj = g.csr()
    ro, ra = orc.relabel(off, adj)
    for k in (3, 4, 5):
        want_c, want_w = orc.clique_walg(ro, ra, k)
        got, st = gat.Plan(g, [clique(k)], "unique").count({"vertex_store": 1}, account=True)
        assert got[0] == want_c
        assert st["alg_bytes"] == want_w, (k, st["alg_bytes"], want_w)


def test_bitmap_walg_matches_oracle(gat, orc, gpu):
    g = gat.Graph.rmat(9, 8, 0.57, 0.19, 0.19, 3)
    off, adj = g.csr()
    ro, ra = orc.relabel(off, adj)
    for kmax in (3, 4, 5, 6):
        want_c, want_w = orc.clique_bitmap_walg(ro, ra, kmax)
        assert want_c == orc.unique_count(off, adj, clique(kmax))
        pats = [clique(k) for k in range(3, kmax + 1)]
        for cap in (256, 32):  # 32: most edges go through the second (mid) tier
            got, st = gat.Plan(g, pats, "unique").count({"vertex_store": 2, "slot_capacity": cap}, account=True)
            if cap == 256:
                assert st["overflow_tasks"] == 0
            assert got[-1] == want_c
            assert st["alg_bytes"] == want_w, (kmax, cap, st["alg_bytes"], want_w)


@pytest.mark.parametrize("gname", ["er200", "rmat8", "cl300"])
def test_bitmap_clique_store(gat, orc, small_graphs, gname):
    """Bitmap rows (vertex_store 2) against the array store and the oracle;
    slot_capacity 32 caps d = |N+(v0)| at 32 in the main kernel so larger
    vertices take the second tier."""
    g = small_graphs[gname]
    off, adj = g.csr()
    ks = [3, 4, 5, 6]
    want = [orc.unique_count(off, adj, clique(k)) for k in ks]
    plan = gat.Plan(g, [clique(k) for k in ks], "unique")
    for cfg in ({"vertex_store": 2}, {"vertex_store": 2, "slot_capacity": 32, "queue_chunk": 1},
                {"vertex_store": 2, "gallop_ratio": 1, "search_linear_threshold": 8}, {"vertex_store": 1}):
        say("bitmap", gname, cfg)
        got, st = plan.count(cfg)
        assert got == want, (cfg, got, want)
    for sub in ([4], [5], [3, 5], [4, 6]):
        got, _ = gat.Plan(g, [clique(k) for k in sub], "unique").count({"vertex_store": 2})
        assert got == [want[ks.index(k)] for k in sub]


def test_bitmap_overflow_and_labeled(gat, orc, gpu):
    # dense graph: many N+(v0) exceed a 32-entry cap. K4 from the oracle; K5
    # from the array store, which the tests above pin to the oracle (the
    # oracle's K5 here takes ~30 s).
    g = gat.Graph.erdos_renyi(60, 0.9, 4)
    off, adj = g.csr()
    arr, _ = gat.Plan(g, [clique(4), clique(5)], "unique").count({"vertex_store": 1})
    want = [orc.unique_count(off, adj, clique(4)), arr[1]]
    assert arr[0] == want[0]
    got, st = gat.Plan(g, [clique(4), clique(5)], "unique").count({"vertex_store": 2, "slot_capacity": 32})
    assert got == want and st["overflow_tasks"] > 0
    lab, _ = gat.Plan(g, [clique(4), clique(5)], "labeled").count({"vertex_store": 2})
    assert lab == [24 * want[0], 120 * want[1]]


def test_walg_and_counts_exact_under_overflow(gat, gpu):
    """A 16-entry LDS set capacity pushes work to the HBM launch (single
    edges when the root has one child, whole tasks otherwise); the counts and
    the W_alg of the step must equal those of a run without overflow."""
    g = gat.Graph.chung_lu(500, 10.0, 2.1, 80, 9)
    # C5 materialises no set (its plan has no LIST op), so nothing overflows
    for pats, overflows in (([HOUSE], True), ([C5], False), (connected_patterns(5), True)):
        plan = gat.Plan(g, pats, "unique")
        c1, s1 = plan.count({"vertex_store": 1}, account=True)
        c2, s2 = plan.count({"vertex_store": 1, "slot_capacity": 16}, account=True)
        assert c1 == c2, pats
        assert (s2["overflow_tasks"] > 0) == overflows, pats
        assert s1["alg_bytes"] == s2["alg_bytes"], (pats, s1["alg_bytes"], s2["alg_bytes"])


@pytest.mark.parametrize("strategy", [0, 1, 2, 3, 4])
def test_intersection_strategies_long_lists(gat, orc, gpu, strategy):
    """Hub-heavy Chung-Lu graph (degrees up to ~2,000): the array store's
    wave-wide intersections of long lists, materialised (K4, K5: the depth-1
    list) and counted (triangles), with every strategy, against the oracle's
    clique counts; blocks of many elements per lane (galloping) and long
    merge runs are exercised. Hub bitmaps off so every probe is a search."""
    g = gat.Graph.chung_lu(20000, 30.0, 2.1, 2000, 13)
    off, adj = g.csr()
    want = orc.clique_counts_mt(off, adj, 5, 8)
    plan = gat.Plan(g, [TRIANGLE, K4, K5], "unique")
    for extra in ({}, {"slot_capacity": 64}):
        cfg = dict({"vertex_store": 1, "intersect_strategy": strategy, "hub_degree": -1}, **extra)
        got, _ = plan.count(cfg)
        say("strategy", strategy, extra, got, want)
        assert got == want, (cfg, got, want)
