"""The three tiers of the bitmap clique store (GPU): the main kernel keeps
the rows of vertices with d = |N+(v0)| <= slot_capacity (at most 256) in
LDS, the second tier d <= 512 in LDS, and the HBM-row kernel takes the rest.
Complete graphs pin every tier by closed forms: on K_m vertex v has
N+(v) = {u < v}, so d takes every value up to m - 1."""
import math

import pytest

from patterns import clique

pytestmark = pytest.mark.gpu


def complete(gat, m):
    src = [i for i in range(m) for j in range(i + 1, m)]
    dst = [j for i in range(m) for j in range(i + 1, m)]
    return gat.Graph.from_edges(m, src, dst)


@pytest.mark.parametrize("cap", [256, 32, 64])
def test_complete_graph_all_tiers(gat, gpu, cap):
    m = 530  # vertices 513..529 (d > 512) reach the HBM-row kernel
    g = complete(gat, m)
    ks = [3, 4, 5]
    got, st = gat.Plan(g, [clique(k) for k in ks], "unique").count({"vertex_store": 2, "slot_capacity": cap})
    assert got == [math.comb(m, k) for k in ks]
    # vertices with d > cap leave the main kernel: v in (cap, m - 1]
    assert st["overflow_tasks"] == m - 1 - cap
    lab, _ = gat.Plan(g, [clique(4)], "labeled").count({"vertex_store": 2, "slot_capacity": cap})
    assert lab == [24 * math.comb(m, 4)]


def test_complete_graph_k6_mid_tier(gat, gpu):
    m = 100  # every |S| fits the first two tiers
    g = complete(gat, m)
    got, st = gat.Plan(g, [clique(6), clique(4)], "unique").count({"vertex_store": 2, "slot_capacity": 32})
    assert got == [math.comb(m, 6), math.comb(m, 4)] and st["overflow_tasks"] > 0
