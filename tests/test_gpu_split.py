"""Split subtasks of the array store (GPU, through the C ABI).

A task that has run longer than split_us once its launch's queue is drained
hands the unfinished rest of its shallowest loop at depth >= 2 to subtask
launches; with split_us = 1 nearly every task does. Counts must equal the
oracle and the unsplit run, W_alg must not change (rebuilt sets are not
charged twice), and a subtask that outgrows LDS must be finished with HBM
buffers, also under shards, reference mode and hipGraph replay.
"""
import sys

import numpy as np
import pytest

from patterns import C5, HOUSE, K4, connected_patterns

pytestmark = pytest.mark.gpu

SPLIT = {"split_us": 1, "vertex_store": 1}
NOSPLIT = {"split_us": -1, "vertex_store": 1}
# 8 waves per CU: the launch shape under which edges overflow in the main
# launch before they split, so that the HBM launch has tasks to hand on
OVERFLOW = {"slot_capacity": 8, "queue_chunk": 1, "waves_per_cu": 8}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def graphs(gat):
    return {"rmat8": gat.Graph.rmat(8, 6, 0.57, 0.19, 0.19, 7),
            "cl300": gat.Graph.chung_lu(300, 8.0, 2.1, 60, 9)}


@pytest.mark.parametrize("gname", ["rmat8", "cl300"])
def test_split_counts_match_oracle(gat, orc, gpu, graphs, gname):
    """K4, house and C5 against the oracle; three seeded 6-vertex patterns
    against the unsplit count (the oracle needs minutes for some of them)."""
    g = graphs[gname]
    off, adj = g.csr()
    rng = np.random.default_rng(3)
    p6 = connected_patterns(6)
    pats = [K4, HOUSE, C5] + [p6[i] for i in rng.choice(len(p6), 3, replace=False)]
    for i, p in enumerate(pats):
        want = orc.unique_count(off, adj, p) if i < 3 else None
        plan = gat.Plan(g, [p], "unique")
        for extra in ({}, OVERFLOW, {"hub_degree": -1}):
            say("split", gname, p, extra)
            base, st0 = plan.count(dict(NOSPLIT, **extra), account=True)
            got, st1 = plan.count(dict(SPLIT, **extra), account=True)
            assert got == base and (want is None or base == [want]), (p, extra)
            assert st1["alg_bytes"] == st0["alg_bytes"], (p, extra)
            assert st0["split_tasks"] == 0


@pytest.mark.parametrize("gname", ["rmat8", "cl300"])
def test_split_budgets_clique_array_store(gat, orc, gpu, graphs, gname):
    """Mid-range budgets make long tasks hand on an ancestor loop first and
    deeper loops after it; every later donation must still record the
    candidate being walked at the ancestor's depth (K5 and K6 were short by
    a few when a closed frame's cursor was taken for that position)."""
    g = graphs[gname]
    off, adj = g.csr()
    ks = [3, 4, 5, 6]
    want = [orc.unique_count(off, adj, "".join("0" if i == j else "1" for i in range(k) for j in range(k)))
            for k in ks]
    plan = gat.Plan(g, ["".join("0" if i == j else "1" for i in range(k) for j in range(k)) for k in ks], "unique")
    for sp in (20, 100, 500, 2000):
        for _ in range(4):
            got, st = plan.count({"vertex_store": 1, "split_us": sp})
            assert got == want, (sp, got, want, st["split_tasks"])


def test_split_path_is_taken(gat, gpu, graphs):
    _, st = gat.Plan(graphs["cl300"], [C5], "unique").count(SPLIT)
    assert st["split_tasks"] > 0 and st["split_overflow"] == 0  # C5 builds no list, so nothing overflows
    # 6-vertex patterns whose first intersection list is built at depth 2:
    # a split subtask can outgrow an 8-entry buffer there and go to the HBM
    # launch
    moved = 0
    for p in ("001000000110100001010001010001001110", "000110001110010001110001110001001110"):
        plan = gat.Plan(graphs["cl300"], [p], "unique")
        want = plan.count(NOSPLIT)[0]
        got, st = plan.count(dict(SPLIT, **OVERFLOW))
        assert got == want and st["split_tasks"] > 0, p
        moved += st["split_overflow"]
        lv = st["split_level"]
        say("split levels", p, lv, st["split_overflow"])
        # the HBM launch (overflowed edges) hands subtasks to the HBM list
        # beyond the copies of LDS-overflowed subtasks
        assert lv[0] > 0, (p, lv)
        assert lv[3] - st["split_overflow"] > 0, (p, lv, st["split_overflow"])
    assert moved > 0


def test_split_reaches_every_lds_level(gat, orc, gpu, graphs):
    """Only loops at depth >= 2 that are not the leaf level split: K6 on the
    array store has two (depths 2 and 3), K7 three, so with split_us = 1
    level-1 subtasks split into level 2 and, for K7, level-2 subtasks into
    level 3; counts equal the oracle's."""
    total = [0, 0, 0, 0]
    for gname in ("rmat8", "cl300"):
        g = graphs[gname]
        off, adj = g.csr()
        for k in (6, 7):
            p = "".join("0" if i == j else "1" for i in range(k) for j in range(k))
            plan = gat.Plan(g, [p], "unique")
            got, st = plan.count(SPLIT)
            say("split levels", gname, "K%d" % k, st["split_level"], st["split_tasks"])
            assert got == [orc.unique_count(off, adj, p)], (gname, k)
            total = [a + b for a, b in zip(total, st["split_level"])]
    assert total[0] > 0 and total[1] > 0 and total[2] > 0, total


def test_split_batched_reference_shards_replay(gat, gpu, graphs):
    g = graphs["rmat8"]
    n = g.v_cnt()
    pats = connected_patterns(5)
    batched = gat.Plan(g, pats, "unique")
    want = batched.count(NOSPLIT)[0]
    assert batched.count(SPLIT)[0] == want
    assert batched.count(dict(SPLIT, **OVERFLOW))[0] == want
    cuts = [0, 3, n // 4, n]
    for lo, hi in zip(cuts, cuts[1:]):
        assert batched.count_shard(lo, hi, SPLIT)[0] == batched.count_shard(lo, hi, NOSPLIT)[0], (lo, hi)
    for p in (K4, HOUSE, C5):
        ref = gat.Plan(g, [p], "reference")
        assert ref.count(SPLIT)[0] == ref.count(NOSPLIT)[0], p
    plan = gat.Plan(g, [HOUSE], "unique")
    eager = plan.count(NOSPLIT)[0]
    for _ in range(3):
        assert plan.count(SPLIT, graph_replays=3)[0] == eager
        assert plan.count(NOSPLIT)[0] == eager
