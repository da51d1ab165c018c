"""Hub bitmaps (GPU): membership tests in high-degree neighbour lists go
through dense bitmaps instead of binary search. Counts must not change with
the hub threshold, in every mode (the reference mode runs on the original
vertex ids, where hubs are not a prefix and go through the hub map), in the
LDS and the HBM-overflow launches, against the oracle."""
import pytest

from patterns import C5, HOUSE, K4, TRIANGLE, connected_patterns

pytestmark = pytest.mark.gpu

THRESHOLDS = (-1, 1, 6, 40, 100000)


def test_hub_thresholds_unique_and_labeled(gat, orc, gpu):
    g = gat.Graph.chung_lu(500, 10.0, 2.1, 80, 9)
    off, adj = g.csr()
    pats = [TRIANGLE, K4, HOUSE, C5] + connected_patterns(5)[2::6]
    want = [orc.unique_count(off, adj, p) for p in pats]
    for thr in THRESHOLDS:
        for extra in ({}, {"slot_capacity": 16}, {"intersect_strategy": 2}, {"gallop_ratio": 4}):
            cfg = dict(extra, hub_degree=thr, vertex_store=1)
            got, _ = gat.Plan(g, pats, "unique").count(cfg)
            assert got == want, (thr, extra)
        lab, _ = gat.Plan(g, [HOUSE], "labeled").count({"hub_degree": thr})
        assert lab == [2 * want[2]], thr


def test_hub_thresholds_reference_mode(gat, orc, gpu):
    from graphautotune_amd._lib import ScheduleData
    g = gat.Graph.rmat(8, 8, 0.57, 0.19, 0.19, 13)
    off, adj = g.csr()
    for p in (HOUSE, C5, K4):
        ref, ok = orc.reference_count(off, adj, orc.schedule(p, ScheduleData))
        assert ok
        for thr in THRESHOLDS:
            for extra in ({}, {"slot_capacity": 16}):
                got, _ = gat.Plan(g, [p], "reference").count(dict(extra, hub_degree=thr))
                assert got == [ref], (p, thr, extra)


def test_hub_threshold_change_rebuilds(gat, gpu):
    g = gat.Graph.rmat(11, 16, 0.57, 0.19, 0.19, 2)
    plan = gat.Plan(g, [HOUSE], "unique")
    base, _ = plan.count({"hub_degree": -1})
    for thr in (32, 8, 32, 1000, 1):
        assert plan.count({"hub_degree": thr})[0] == base, thr
