"""Shards of a count on the GPU (gat_count_shard through the C ABI).

A shard counts the embeddings whose Schedule vertex 0 maps to a root in
[lo, hi) and returns raw sums (divisor * count over a partition).

Small graphs: every shard equals the oracle's restricted count under the
plan's ordered pattern and restrictions, and a partition adds up to
divisor * gat_count. This covers cliques in both stores, house, C5,
5-vertex patterns whose divisor exceeds 1 and a 6-vertex pattern, under the
default config and one that sends most tasks to the HBM overflow launch.

Real sizes (the benchmark configurations; shards sized so a launch takes
seconds):
* R-MAT scale 20, K4 + K5: single-root shards against the oracle's
  per-top-vertex bitmap clique count; array store equal to bitmap store.
* Chung-Lu 4,194,304 vertices (mean degree 24, exponent 2.1, cap 50,000,
  seed 21): C5 on the hub core against the oracle (on these roots every
  other cycle vertex lies below the root; the 16 largest hubs form a clique,
  so roots [0, 16) hold C(16,5) * 12 cycles); house and C5 batched equal
  plain; shards add up; overflow-heavy and hub-free configs agree.
* R-MAT scale 22 with all 21 connected 5-vertex and 10 seeded 6-vertex
  patterns: batched equals plain per shard, and shards add up.
"""
import sys

import numpy as np
import pytest

from patterns import C5, HOUSE, K4, K5, TRIANGLE, connected_patterns

pytestmark = pytest.mark.gpu

OVERFLOW = {"slot_capacity": 8, "queue_chunk": 1}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ordered(plan, i):
    d = plan.schedule(i).data
    return list(d.adj_rows[:d.v_cnt]), plan.count_info(i)


def shard_cuts(n):
    return [0, 1, n // 9, n // 3, n - 1, n]


def small_patterns(gat, g):
    """Cliques, house, C5, three 5-vertex patterns with divisor > 1, one
    6-vertex pattern."""
    multi = [p for p in connected_patterns(5) if gat.Plan(g, [p], "unique").count_info(0)["divisor"] > 1]
    assert len(multi) >= 3
    return [TRIANGLE, K4, K5, HOUSE, C5] + multi[:3] + [connected_patterns(6)[40]]


@pytest.mark.parametrize("gname", ["rmat8", "cl300"])
def test_shards_match_oracle(gat, orc, gpu, gname):
    g = (gat.Graph.rmat(8, 6, 0.57, 0.19, 0.19, 3) if gname == "rmat8"
         else gat.Graph.chung_lu(300, 6.0, 2.1, 40, 5))
    off, adj = g.csr()
    roff, radj = orc.relabel(off, adj)
    cuts = shard_cuts(g.v_cnt())
    for p in small_patterns(gat, g):
        plan = gat.Plan(g, [p], "unique")
        rows, info = ordered(plan, 0)
        want = [orc.restricted_count(roff, radj, rows, info["restrictions"], lo, hi)
                for lo, hi in zip(cuts, cuts[1:])]
        total = plan.count()[0][0]
        assert sum(want) == info["divisor"] * total, p
        cfgs = [{}, OVERFLOW]
        if p in (TRIANGLE, K4, K5):
            cfgs += [{"vertex_store": 1}, {"vertex_store": 2, "slot_capacity": 32}]
        for cfg in cfgs:
            say("shards", gname, p, cfg)
            got = [plan.count_shard(lo, hi, cfg)[0][0] for lo, hi in zip(cuts, cuts[1:])]
            assert got == want, (p, cfg)
        lab = gat.Plan(g, [p], "labeled")
        assert [lab.count_shard(lo, hi)[0][0] for lo, hi in zip(cuts, cuts[1:])] == want


def test_batched_and_reference_shards(gat, gpu):
    g = gat.Graph.rmat(8, 6, 0.57, 0.19, 0.19, 4)
    n = g.v_cnt()
    cuts = shard_cuts(n)
    pats = connected_patterns(5)
    batched = gat.Plan(g, pats, "unique")
    plains = [gat.Plan(g, [p], "unique") for p in pats]
    for lo, hi in zip(cuts, cuts[1:]):
        say("batched shard", lo, hi)
        assert batched.count_shard(lo, hi)[0] == [pl.count_shard(lo, hi)[0][0] for pl in plains]
    assert batched.count_shard(5, 5)[0] == [0] * len(pats)
    # reference mode reports raw sums itself: a partition adds up to gat_count
    for p in (K4, HOUSE):
        ref = gat.Plan(g, [p], "reference")
        assert sum(ref.count_shard(lo, hi)[0][0] for lo, hi in zip(cuts, cuts[1:])) == ref.count()[0][0]


def top_vertex_cliques(orc, roff, radj, k, v):
    n = len(roff) - 1
    return orc.clique_bitmap_walg_sample(roff, radj, k, n, v)[0]


def test_rmat20_clique_shards(gat, orc, gpu):
    g = gat.Graph.rmat(20, 16, 0.57, 0.19, 0.19, 20)
    off, adj = g.csr()
    roff, radj = orc.relabel(off, adj)
    n = g.v_cnt()
    plan = gat.Plan(g, [K4, K5], "unique")
    for v in (0, 9, 300, 2000, 20000, 200000, n - 1000):
        say("rmat20 root", v)
        got = plan.count_shard(v, v + 1)[0]
        assert got == [top_vertex_cliques(orc, roff, radj, 4, v), top_vertex_cliques(orc, roff, radj, 5, v)], v
    lo, hi = 3000, 3400
    bitmap = plan.count_shard(lo, hi)[0]
    assert gat.Plan(g, [K4], "unique").count_shard(lo, hi, {"vertex_store": 1})[0][0] == bitmap[0]
    halves = [plan.count_shard(lo, 3100)[0], plan.count_shard(3100, hi)[0]]
    assert [a + b for a, b in zip(*halves)] == bitmap


@pytest.fixture(scope="module")
def chung_lu_full(gat):
    return gat.Graph.chung_lu(4194304, 24.0, 2.1, 50000, 21)


def test_chung_lu_full_c5_hub_core(gat, orc, gpu, chung_lu_full):
    g = chung_lu_full
    off, adj = g.csr()
    roff, radj = orc.relabel(off, adj)
    plan = gat.Plan(g, [C5], "unique")
    rows, info = ordered(plan, 0)
    assert info["divisor"] == 1 and {b for a, b in info["restrictions"] if a == 0} == {1, 2, 3, 4}
    assert plan.count_shard(0, 16)[0] == [4368 * 12]  # C(16,5) * 12: the 16 largest hubs are a clique
    for lo, hi in ((0, 40), (30, 44)):
        say("cl4M C5 shard", lo, hi)
        want = orc.restricted_count(roff, radj, rows, info["restrictions"], lo, hi)
        assert plan.count_shard(lo, hi)[0] == [want]
    k4 = gat.Plan(g, [K4], "unique")
    for v in (3, 500, 40000, 2000000):
        assert k4.count_shard(v, v + 1)[0] == [top_vertex_cliques(orc, roff, radj, 4, v)], v


def test_chung_lu_full_house_c5_properties(gat, gpu, chung_lu_full):
    g = chung_lu_full
    n = g.v_cnt()
    mid = n // 2
    house = gat.Plan(g, [HOUSE], "unique")
    c5 = gat.Plan(g, [C5], "unique")
    both = gat.Plan(g, [HOUSE, C5], "unique")
    say("cl4M batched vs plain")
    assert both.count_shard(mid, mid + 2)[0] == [house.count_shard(mid, mid + 2)[0][0],
                                                 c5.count_shard(mid, mid + 2)[0][0]]
    say("cl4M house shards")
    whole = house.count_shard(mid, mid + 4096)[0][0]
    assert whole > 0
    assert house.count_shard(mid, mid + 1000)[0][0] + house.count_shard(mid + 1000, mid + 4096)[0][0] == whole
    assert house.count_shard(mid, mid + 4096, {"slot_capacity": 16})[0][0] == whole
    part = house.count_shard(mid, mid + 256)[0][0]
    assert house.count_shard(mid, mid + 256, {"hub_degree": -1})[0][0] == part


def test_rmat22_batched_31_shards(gat, orc, gpu):
    """The batched configuration's graph (R-MAT scale 22) and 31 patterns:
    batched equals plain and shards add up on mid-degree roots and on the
    64 lightest roots of degree >= 2; on the 64 largest hubs the oracle's
    restricted count is reached for the patterns whose restrictions make the
    root the largest id of every embedding (the rest of the embedding then
    lies among the hubs; for the other patterns a root's embeddings run
    through every hub neighbourhood, beyond the oracle at this size)."""
    import numpy as np
    from graphautotune_amd import workloads
    g = workloads.graph("rmat22")
    pats = workloads.batched_patterns()
    n = g.v_cnt()
    lo = n // 2
    batched = gat.Plan(g, pats, "unique")
    say("rmat22 batched shard")
    b = batched.count_shard(lo, lo + 4)[0]
    halves = [batched.count_shard(lo, lo + 2)[0], batched.count_shard(lo + 2, lo + 4)[0]]
    assert [x + y for x, y in zip(*halves)] == b
    say("rmat22 plain shards")
    assert b == [gat.Plan(g, [p], "unique").count_shard(lo, lo + 4)[0][0] for p in pats]
    off, adj = g.csr()
    roff, radj = orc.relabel(off, adj)
    deg = np.diff(roff)
    hi = int(np.nonzero(deg >= 2)[0][-1]) + 1
    light = (hi - 64, hi)
    say("rmat22 light roots", light)
    lb = batched.count_shard(*light)[0]
    assert lb == [gat.Plan(g, [p], "unique").count_shard(*light)[0][0] for p in pats]
    assert [x + y for x, y in zip(batched.count_shard(light[0], light[0] + 20)[0],
                                  batched.count_shard(light[0] + 20, light[1])[0])] == lb
    checked = 0
    for i in range(len(pats)):
        plan = gat.Plan(g, [pats[i]], "unique")
        rows, info = ordered(plan, 0)
        if {b_ for a_, b_ in info["restrictions"] if a_ == 0} == set(range(1, len(rows))):
            say("rmat22 hub core oracle", pats[i])
            assert plan.count_shard(0, 64)[0][0] == orc.restricted_count(roff, radj, rows, info["restrictions"],
                                                                         0, 64), pats[i]
            checked += 1
    assert checked >= 2  # K5 and the 5-vertex pattern with restrictions (0,1) .. (0,4)


def test_count_sharded_equals_count(gat, gpu):
    """The sharded driver (many tiny shards) reports what gat_count does, in
    every mode, for patterns whose divisor exceeds 1 too."""
    from graphautotune_amd.sharded import count_sharded, edge_shards
    g = gat.Graph.rmat(9, 8, 0.57, 0.19, 0.19, 5)
    multi = [p for p in connected_patterns(5) if gat.Plan(g, [p], "unique").count_info(0)["divisor"] > 1][:3]
    for pats, mode in (([HOUSE, C5], "unique"), (multi, "unique"), (multi, "labeled"), ([K4, HOUSE], "reference")):
        plan = gat.Plan(g, pats, mode)
        got, shards = count_sharded(plan, target_s=1e-4, first=3)
        assert got == plan.count()[0], (pats, mode)
        # array-store plans are walked in directed-edge shards, clique plans in root shards
        end = int(plan.run_offsets()[-1]) if edge_shards(plan) else g.v_cnt()
        assert len(shards) > 3 and shards[0][0] == 0 and shards[-1][1] == end


def test_reference_mode_shards_match_oracle(gat, orc, gpu):
    """Reference mode (input ids, the reference's restrictions and IEP):
    every shard equals the oracle's restatement of the reference engine over
    the same first vertices."""
    from graphautotune_amd._lib import ScheduleData
    for g in (gat.Graph.rmat(8, 6, 0.57, 0.19, 0.19, 3), gat.Graph.chung_lu(300, 6.0, 2.1, 40, 5)):
        off, adj = g.csr()
        cuts = shard_cuts(g.v_cnt())
        for p in (TRIANGLE, K4, HOUSE, C5):
            s = orc.schedule(p, ScheduleData)
            want = [orc.reference_count_roots(off, adj, s, lo, hi) for lo, hi in zip(cuts, cuts[1:])]
            if not all(ok for _, ok in want):
                continue
            plan = gat.Plan(g, [p], "reference")
            assert not plan.count_info(0)["relabeled"]
            for cfg in ({}, OVERFLOW):
                got = [plan.count_shard(lo, hi, cfg)[0][0] for lo, hi in zip(cuts, cuts[1:])]
                assert got == [v for v, _ in want], (p, cfg)


def test_edge_shards_partition_counts(gat, gpu):
    """gat_count_shard_edges: raw sums over a partition of the directed edges
    of the plan's CSR add up to the root shards' and to divisor x count, in
    unique and reference modes (reference plans run on the input ids), with
    splitting and overflow; a heavy root's edges split across shards; clique
    plans on the bitmap store refuse edge shards, on the array store take
    them."""
    from graphautotune_amd._lib import GAT_EUNSUPPORTED, GatError
    g = gat.Graph.chung_lu(400, 10.0, 2.1, 80, 3)
    p6 = connected_patterns(6)[40]
    for pats, mode in (([HOUSE, C5, p6], "unique"), ([HOUSE, K4], "reference")):
        plan = gat.Plan(g, pats, mode)
        off = plan.run_offsets()
        m = int(off[-1])
        whole = plan.count_shard(0, g.v_cnt())[0]
        for cfg in ({}, OVERFLOW, {"split_us": 1, "vertex_store": 1}):
            cuts = [0, 1, int(off[1]) // 2, int(off[1]), m // 3, m - 1, m]
            parts = [plan.count_shard_edges(a, b, cfg)[0] for a, b in zip(cuts, cuts[1:])]
            assert [sum(x) for x in zip(*parts)] == whole, (pats, mode, cfg)
        if mode == "unique":
            counts = plan.count()[0]
            infos = [plan.count_info(i) for i in range(len(pats))]
            assert [w // i["divisor"] * i["multiplier"] for w, i in zip(whole, infos)] == counts
    clique = gat.Plan(g, [K4], "unique")
    with pytest.raises(GatError) as e:
        clique.count_shard_edges(0, 10)
    assert e.value.code == GAT_EUNSUPPORTED
    m = int(clique.run_offsets()[-1])
    halves = [clique.count_shard_edges(0, m // 2, {"vertex_store": 1})[0][0],
              clique.count_shard_edges(m // 2, m, {"vertex_store": 1})[0][0]]
    assert sum(halves) == clique.count_shard(0, g.v_cnt())[0][0]
