"""The benchmark configurations at checkable sizes (GPU, through the C ABI).

* Erdos-Renyi G(10000, 16/9999), seed 1: triangles against the oracle, and
  reference = 3x, labeled = 6x the unique count.
* R-MAT scale 20 (edge factor 16, .57/.19/.19, seed 20): K4 from the bitmap
  store equals the array store; K4 and K5 batched equal the single-pattern
  runs (size-independent properties at full size).
* Chung-Lu (gamma 2.1), scaled down to 600 vertices so the oracle finishes:
  house and C5, unique against the oracle, reference against the oracle's
  restatement.
* Batched 21 five-vertex + 10 seeded random six-vertex patterns on R-MAT
  scale 9: one launch equals the 31 plain runs, and a sample equals the
  oracle.
* The pm CLI: Answer line and the tuner's result files.
"""
import os
import subprocess

import numpy as np
import pytest

from patterns import C5, HOUSE, K4, K5, TRIANGLE, connected_patterns

pytestmark = pytest.mark.gpu


def test_er10k_triangles(gat, orc, gpu):
    g = gat.Graph.erdos_renyi(10000, 16 / 9999, 1)
    off, adj = g.csr()
    want = orc.unique_count(off, adj, TRIANGLE)
    assert gat.Plan(g, [TRIANGLE], "unique").count()[0] == [want]
    assert gat.Plan(g, [TRIANGLE], "reference").count()[0] == [3 * want]
    assert gat.Plan(g, [TRIANGLE], "labeled").count()[0] == [6 * want]
    assert gat.Plan(g, [TRIANGLE], "unique").count({"vertex_store": 1})[0] == [want]


def test_rmat20_clique_properties(gat, gpu):
    g = gat.Graph.rmat(20, 16, 0.57, 0.19, 0.19, 20)
    k4_bitmap, _ = gat.Plan(g, [K4], "unique").count({"vertex_store": 2})
    k4_array, _ = gat.Plan(g, [K4], "unique").count({"vertex_store": 1})
    assert k4_bitmap == k4_array
    both, _ = gat.Plan(g, [K4, K5], "unique").count()
    k5, _ = gat.Plan(g, [K5], "unique").count()
    assert both == [k4_bitmap[0], k5[0]]


@pytest.mark.parametrize("scale", [12, 16, 20])
def test_rmat_clique_counts_golden(gat, gpu, scale):
    """Whole-graph K3, K4 and K5 of the benchmark graphs equal the oracle's
    counts committed in tests/golden/clique_counts.json, batched and in the
    tuned configuration of the benchmark."""
    import json
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "clique_counts.json")) as f:
        want = json.load(f)["rmat%d" % scale]
    g = gat.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, 20)
    assert (g.v_cnt(), g.undirected_edges()) == (want["vertices"], want["undirected_edges"])
    with open(os.path.join(os.path.dirname(here), "graphautotune_amd", "tuning", "tuned_configs.json")) as f:
        tuned = json.load(f)["rmat20-k4k5"]["config"]
    plan = gat.Plan(g, [TRIANGLE, K4, K5], "unique")
    for cfg in ({}, tuned):
        got, _ = plan.count(cfg)
        print("rmat%d" % scale, cfg, got, flush=True)
        assert got == [want["K3"], want["K4"], want["K5"]], cfg


def test_chung_lu_house_c5(gat, orc, gpu):
    from graphautotune_amd._lib import ScheduleData
    g = gat.Graph.chung_lu(600, 12.0, 2.1, 40, 21)
    off, adj = g.csr()
    for p in (HOUSE, C5):
        want = orc.unique_count(off, adj, p)
        for cfg in ({}, {"slot_capacity": 16}, {"gallop_ratio": 4}):
            assert gat.Plan(g, [p], "unique").count(cfg)[0] == [want], (p, cfg)
        ref, ok = orc.reference_count(off, adj, orc.schedule(p, ScheduleData))
        assert ok and gat.Plan(g, [p], "reference").count()[0] == [ref]
    both, _ = gat.Plan(g, [HOUSE, C5], "unique").count()
    assert both == [orc.unique_count(off, adj, HOUSE), orc.unique_count(off, adj, C5)]


def test_batched_21_plus_10(gat, orc, gpu):
    from graphautotune_amd import workloads
    pats = workloads.batched_patterns()
    g = gat.Graph.rmat(9, 16, 0.57, 0.19, 0.19, 22)
    batched, _ = gat.Plan(g, pats, "unique").count()
    plain = [gat.Plan(g, [p], "unique").count()[0][0] for p in pats]
    assert batched == plain
    small = gat.Graph.rmat(6, 8, 0.57, 0.19, 0.19, 22)
    off, adj = small.csr()
    got, _ = gat.Plan(small, pats, "unique").count()
    for i in range(0, len(pats), 3):
        assert got[i] == orc.unique_count(off, adj, pats[i]), pats[i]


def test_pm_cli(gat, gpu, tmp_path):
    g = gat.Graph.rmat(10, 8, 0.57, 0.19, 0.19, 3)
    g.save(tmp_path / "g.bin")
    pm = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "graphautotune_amd", "bin", "pm")
    r = subprocess.run([pm, "g.bin", K4, "abc123", "--root", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ans = int([l for l in r.stdout.splitlines() if l.startswith("Answer:")][0].split()[1])
    uniq = gat.Plan(g, [K4], "unique").count()[0][0]
    assert ans == 4 * uniq  # reference semantics by default
    res = tmp_path / "tuning_results" / "abc123"
    assert int((res / "count.txt").read_text()) == ans
    assert float((res / "time.txt").read_text()) > 0
    r2 = subprocess.run([pm, "g.bin", K4, "--root", str(tmp_path), "--mode", "unique", "--config",
                         "queue_chunk=4,slot_capacity=256"], capture_output=True, text=True)
    assert r2.returncode == 0 and ("Answer: %d" % uniq) in r2.stdout
