"""Tuner host logic (CPU only): config space, fingerprints, simulated
annealing, grid search and the driver's infeasible-config mapping, with a
stub engine and a stub cost model."""
import hashlib
import random

import numpy as np
import pytest

from graphautotune_amd.tuning import space as sp
from graphautotune_amd.tuning.driver import FLOAT_INF
from graphautotune_amd.tuning.manipulator import Manipulator
from graphautotune_amd.tuning.tuner import Tuner, grid_search


def test_space_matches_reference_values():
    # tuning/tuning/config/details.py:11-13 (SetSearchType) kept verbatim
    assert sp.BY_NAME["search_linear_threshold"].values == [0, 4, 16, 64, 2 ** 31 - 1]
    full = sp.ConfigSpace()
    assert full.size() == 3 * 4 * 4 * 6 * 5 * 3 * 5 * 2  # intersect_strategy: 5 values
    assert len(list(sp.ConfigSpace(["block_waves", "queue_chunk"]).all())) == 12


def test_fingerprint_is_md5_of_sorted_index_vector():
    # base.py:218-222,259-262: indices of the values, keys sorted by name
    c = sp.Config({"queue_chunk": 64, "block_waves": 2})
    idx = [sp.BY_NAME[k].values.index(c.values[k]) for k in sorted(c.values)]
    assert c.get_list() == idx
    assert c.fingerprint() == hashlib.md5(",".join(map(str, idx)).encode()).hexdigest()
    with pytest.raises(ValueError):
        sp.Config({"queue_chunk": 3})


def test_random_walk_changes_exactly_one_knob():
    s = sp.ConfigSpace()
    rng = random.Random(1)
    c = s.random(rng)
    for _ in range(50):
        d = s.random_walk(c, rng)
        assert sum(c.values[k] != d.values[k] for k in c.values) == 1
        c = d


class StubDriver:
    """Runtime surface with a known optimum; some configs infeasible."""

    def __init__(self):
        self.calls = 0

    def cost(self, c):
        v = c.values
        if v["block_waves"] == 4 and v["slot_capacity"] == 1024:
            return FLOAT_INF  # e.g. too much LDS
        return (1.0 + abs(v["waves_per_cu"] - 12) / 4 + abs(np.log2(v["queue_chunk"]) - 2) / 2
                + (v["slot_capacity"] != 256) * 0.5 + (v["block_waves"] != 1) * 0.25)

    def run(self, c):
        self.calls += 1
        return self.cost(c)

    def batch_run(self, cs):
        return [self.run(c) for c in cs]


class StubModel:
    """Perfect-knowledge cost model: checks the search logic, not learning."""

    def __init__(self, drv):
        self.drv, self.xs, self.ys = drv, [], []

    def predict_list(self, cs):
        return np.array([min(self.drv.cost(c), 1e9) for c in cs])

    def is_config_exist(self, c):
        return c.fingerprint() in {x.fingerprint() for x in self.xs}

    def update_list(self, xs, ys):
        self.xs += xs
        self.ys += ys

    def get_best_config(self):
        i = int(np.argmin(self.ys))
        return self.xs[i], self.ys[i]


def test_sa_tuner_finds_optimum_with_a_good_model():
    space = sp.ConfigSpace(["block_waves", "waves_per_cu", "queue_chunk", "slot_capacity"])
    drv = StubDriver()
    m = Manipulator(StubModel(drv), drv, space, seed=3)
    best, t = Tuner(m, max_round=6, batch_size=5, warmup_examples=8).tune()
    assert t == pytest.approx(1.0)
    assert best.values["waves_per_cu"] == 12 and best.values["queue_chunk"] == 4
    assert drv.calls <= 8 + 6 * 5  # every pick is measured, nothing more


def test_grid_search_is_the_judge():
    space = sp.ConfigSpace(["block_waves", "waves_per_cu", "queue_chunk", "slot_capacity"])
    drv = StubDriver()
    best, t, results = grid_search(drv, space)
    assert len(results) == space.size()
    assert t == pytest.approx(1.0)
    assert sum(1 for _, y in results if y == FLOAT_INF) == 4 * 4  # infeasible corner skipped, not fatal


def test_manipulator_drops_infeasible_runs():
    space = sp.ConfigSpace(["block_waves", "slot_capacity"])
    drv = StubDriver()
    m = Manipulator(StubModel(drv), drv, space)
    bad = sp.Config({"block_waves": 4, "slot_capacity": 1024})
    xs, ys = m.run_batch([bad, sp.Config()])
    assert xs == [sp.Config()] and ys[0] < FLOAT_INF


def test_workload_patterns():
    """The product's pattern enumeration equals the committed fixture (and
    OEIS A001349: 2, 6, 21, 112 connected graphs on 3..6 vertices); the
    batched set is the 21 five-vertex patterns plus 10 distinct six-vertex
    patterns of edge density 0.4-0.8."""
    from graphautotune_amd import workloads
    from patterns import connected_patterns
    for n, count in ((3, 2), (4, 6), (5, 21), (6, 112)):
        assert list(workloads.connected_patterns(n)) == connected_patterns(n)
        assert len(connected_patterns(n)) == count
    b = workloads.batched_patterns()
    assert len(b) == 31 == len(set(b)) and b[:21] == connected_patterns(5)
    assert all(len(p) == 36 and 0.4 <= workloads.edge_density(p) <= 0.8 for p in b[21:])
    assert workloads.batched_patterns() == b  # seeded


def test_config_study_helpers():
    """The config-ranking study's no-model baseline is the exact expectation
    of the best of k random picks (checked against every k-subset), and its
    space drops the always-linear hub scans and pins the array store."""
    import itertools

    import numpy as np

    from graphautotune_amd.tuning import config_study as cs
    rng = np.random.default_rng(5)
    for n in (1, 4, 7):
        v = rng.uniform(1, 10, n)
        for k in range(1, n + 1):
            brute = np.mean([min(c) for c in itertools.combinations(v, k)])
            assert cs.expected_min_of_random(v, k) == pytest.approx(brute)
    assert cs.expected_min_of_random([3.0, 1.0], 5) == pytest.approx(1.0)
    space = cs.space()
    names = {p.name: p for p in space.params}
    assert 64 not in names["search_linear_threshold"].values and "vertex_store" not in names
    assert all(c.values["vertex_store"] == 1 for c in space.random_configs(20))


def test_study_shard_picking_and_probe(monkeypatch):
    """pick_shard grows a root shard from one root until the default config
    takes the target, moves to lighter roots when one root is too heavy and
    to heavier ones when all roots from lo on are too light; the Driver's
    probe scores a config whose first eighth already takes its limit by
    that time scaled to the shard, without the full run."""
    import numpy as np

    from graphautotune_amd.tuning import config_study as cs
    from graphautotune_amd.tuning import driver as drv_mod

    n = 1 << 16
    # root r costs cost[r] ms: heavy below n/16 and at n/16 itself, light far out
    cost = np.where(np.arange(n) < n // 8, 200.0, 0.05)
    calls = []

    class StubPlan:
        def __init__(self, graph, patterns, mode="unique", **kw):
            self.scale = kw.get("scale", 1.0)

        def count_shard(self, lo, hi, cfg=None):
            calls.append((lo, hi, dict(cfg or {})))
            slow = 50.0 if (cfg or {}).get("queue_chunk") == 64 else 1.0
            return [hi - lo], {"seconds": slow * float(cost[lo:hi].sum()) / 1e3}

    graph = type("G", (), {"v_cnt": lambda self: n})()
    monkeypatch.setattr(cs.core, "Plan", StubPlan)
    lo, hi, t = cs.pick_shard(graph, "p", target_ms=40.0)
    assert lo == n // 4 and t * 1e3 >= 40.0 and t * 1e3 <= 80.0 + 1e-9  # one root at n/16 took 200 ms
    assert (hi - lo) * 0.05 == pytest.approx(t * 1e3)
    cost[:] = 0.0001  # every root light: all roots from n/16 on stay under target/4 ...
    lo2, hi2, t2 = cs.pick_shard(graph, "p", target_ms=40.0)
    assert lo2 < n // 16 and hi2 == n  # ... so the shard moved to heavier (lower) roots

    monkeypatch.setattr(drv_mod.core, "Plan", StubPlan)
    cost[:] = 0.1
    d = drv_mod.Driver(graph, ["p"], reps=1, warmup=0, roots=(0, 800))
    td = d.run({})
    d.probe = (100, 2.5 * td)
    calls.clear()
    t_slow = d.run({"queue_chunk": 64})  # 50x slower: the first 100 roots take 0.5 s > 2.5 x 80 ms
    assert calls == [(0, 100, {"queue_chunk": 64})] and t_slow == pytest.approx(50 * 0.1 * 800 / 1e3)
    calls.clear()
    assert d.run({"queue_chunk": 4}) == pytest.approx(td) and len(calls) == 2  # probe, then the full shard
