"""Shards of a count, host side (CPU only): what gat_plan_count_info reports
and the oracle's restricted count. Over a partition of the roots, the
oracle's restricted counts under a plan's ordered pattern and restriction
set add up to divisor * unique count (labeled count / |Aut|), for every
connected 3- to 5-vertex pattern in unique and labeled modes; the argument
checks of gat_count_shard run before any device work."""
import pytest

from patterns import K4, K5, TRIANGLE, connected_patterns


def plan_pattern(plan, i):
    info = plan.count_info(i)
    k = plan.schedule(i).data.v_cnt
    return list(plan.schedule(i).data.adj_rows[:k]), info


@pytest.mark.parametrize("k", [3, 4, 5])
def test_shard_partition_sums_to_divisor_times_unique(gat, orc, k):
    g = gat.Graph.rmat(6, 8, 0.57, 0.19, 0.19, 3)
    off, adj = g.csr()
    roff, radj = orc.relabel(off, adj)
    n = len(off) - 1
    cuts = [0, n // 7, n // 2, n - 3, n]
    for p in connected_patterns(k):
        unique = orc.unique_count(off, adj, p)
        for mode in ("unique", "labeled"):
            plan = gat.Plan(g, [p], mode)
            rows, info = plan_pattern(plan, 0)
            assert info["relabeled"]
            assert info["multiplier"] == (1 if mode == "unique" else orc.automorphisms(p))
            parts = [orc.restricted_count(roff, radj, rows, info["restrictions"], cuts[i], cuts[i + 1])
                     for i in range(len(cuts) - 1)]
            assert sum(parts) == info["divisor"] * unique, (p, mode, parts)


def test_clique_roots_are_the_top_vertex(gat, orc):
    """For cliques every restriction chain ends at vertex 0: psi(0) is the
    clique's highest id, so a shard holds the cliques whose top vertex it
    holds (what the bitmap store counts per vertex)."""
    g = gat.Graph.rmat(6, 8, 0.57, 0.19, 0.19, 4)
    off, adj = g.csr()
    roff, radj = orc.relabel(off, adj)
    n = len(off) - 1
    for p, kk in ((TRIANGLE, 3), (K4, 4), (K5, 5)):
        plan = gat.Plan(g, [p], "unique")
        rows, info = plan_pattern(plan, 0)
        assert info["divisor"] == 1
        for v in range(0, n, 5):
            want = orc.clique_bitmap_walg_sample(roff, radj, kk, n, v)[0]
            assert orc.restricted_count(roff, radj, rows, info["restrictions"], v, v + 1) == want


def test_shard_argument_checks(gat):
    from graphautotune_amd._lib import GAT_EINVAL, GatError
    g = gat.Graph.erdos_renyi(40, 0.2, 1)
    plan = gat.Plan(g, [K4], "unique")
    for lo, hi in ((-1, 3), (5, 4), (0, 41)):
        with pytest.raises(GatError) as e:
            plan.count_shard(lo, hi)
        assert e.value.code == GAT_EINVAL
    with pytest.raises(GatError):
        plan.count_info(1)


class StubPlan:
    """Stands in for a Plan: shard raw sums are sums over roots of known
    per-root values, and each shard 'takes' time proportional to its work."""

    class G:
        def __init__(self, n):
            self.n = n

        def v_cnt(self):
            return self.n

    def __init__(self, per_root, divisor=2, multiplier=3):
        self.per_root = per_root
        self.graph = self.G(len(per_root))
        self.patterns = ["p"]
        self.divisor, self.multiplier = divisor, multiplier
        self.calls = []

    def count_shard(self, lo, hi, config=None):
        self.calls.append((lo, hi))
        work = sum(self.per_root[lo:hi])
        return [work], {"seconds": 1e-3 * (hi - lo) + 1e-6 * work, "split_tasks": 0}

    def count_shard_edges(self, lo, hi, config=None):  # a bitmap-store (clique) plan: root shards only
        from graphautotune_amd._lib import GAT_EUNSUPPORTED, GatError
        raise GatError(GAT_EUNSUPPORTED, "edge shards need the array store")

    def count_info(self, i):
        return {"divisor": self.divisor, "multiplier": self.multiplier, "relabeled": True, "restrictions": []}


def test_sharded_driver_partitions_roots_and_divides_once():
    from graphautotune_amd.sharded import count_sharded, raw_sharded
    per_root = [4000] * 5 + [2] * 3000 + [0] * 17
    plan = StubPlan(per_root)
    counts, shards = count_sharded(plan, target_s=0.05)
    assert counts == [sum(per_root) // 2 * 3]
    # the shards partition [0, n) in order and adapt their size
    assert shards[0][0] == 0 and shards[-1][1] == len(per_root)
    assert all(a[1] == b[0] for a, b in zip(shards, shards[1:]))
    assert max(hi - lo for lo, hi, _ in shards) > 10 * (shards[0][1] - shards[0][0])
    raw, sh = raw_sharded(plan, target_s=0.05, roots=(100, 2000))
    assert raw == [sum(per_root[100:2000])] and sh[0][0] == 100 and sh[-1][1] == 2000
    with pytest.raises(RuntimeError):
        count_sharded(StubPlan([1, 2, 4]), target_s=1.0)  # raw 7 is not a multiple of 2


class _StubPlan:
    """count_shard over synthetic per-root raw sums and costs (host only):
    root v holds value[v] embeddings and costs cost[v] seconds, a launch a
    fixed 1 ms on top."""

    def __init__(self, values, costs):
        import numpy as np
        self.values, self.costs = np.asarray(values, dtype=np.float64), np.asarray(costs)
        self.patterns = ["p"]
        n = len(values)
        self.graph = type("G", (), {"v_cnt": lambda _self: n})()
        self.calls = 0

    def count_shard(self, lo, hi, cfg=None):
        self.calls += 1
        return [int(self.values[lo:hi].sum())], {"seconds": 1e-3 + float(self.costs[lo:hi].sum()), "split_tasks": 0}

    def count_shard_edges(self, lo, hi, config=None):
        from graphautotune_amd._lib import GAT_EUNSUPPORTED, GatError
        raise GatError(GAT_EUNSUPPORTED, "edge shards need the array store")


def test_stratified_estimate_exact_and_unbiased():
    """Strata of doubling root ranges: with every block sampled the estimate
    is the exact total with zero error; with a few blocks per stratum the
    mean over seeds is within a few standard errors of the total; block
    sizes grow where roots are cheap (a block lasts about min_block_s)."""
    import numpy as np

    from graphautotune_amd.sharded import stratified_estimate
    rng = np.random.default_rng(0)
    n = 5000
    # heavy, skewed roots first (degree order), light ones after
    costs = 2.0 / (1 + np.arange(n)) ** 1.2 + rng.uniform(0, 2e-4, n)
    values = np.floor(1e6 * costs * rng.uniform(0.5, 1.5, n))
    exact = stratified_estimate(_StubPlan(values, costs), per_stratum=10 ** 6, min_block_s=0.01)
    assert exact["raw"][0] == pytest.approx(values.sum()) and exact["raw_se"][0] == 0.0
    assert exact["seconds"] == pytest.approx(costs.sum(), rel=0.2)
    ests, ses = [], []
    for seed in range(40):
        plan = _StubPlan(values, costs)
        e = stratified_estimate(plan, per_stratum=4, min_block_s=0.01, seed=seed)
        ests.append(e["raw"][0])
        ses.append(e["raw_se"][0])
    assert abs(np.mean(ests) - values.sum()) < 3 * np.mean(ses) / np.sqrt(len(ests)) + 1e-9 * values.sum()
    last = e["strata"][-1]
    assert last["block"] > 1 and last["sampled"] == 4


class _EdgeStubPlan:
    """An array-store plan over a CSR of known per-edge raw values: edge
    shards sum them; root shards are edge shards over the roots' edges."""

    def __init__(self, degrees, seed=1):
        import numpy as np
        rng = np.random.default_rng(seed)
        self.off = np.concatenate([[0], np.cumsum(degrees)]).astype(np.int64)
        self.per_edge = rng.integers(0, 1000, int(self.off[-1]))
        self.patterns = ["p"]
        n = len(degrees)
        self.graph = type("G", (), {"v_cnt": lambda _self: n})()
        self.edge_calls = []

    def run_offsets(self):
        return self.off

    def count_shard_edges(self, lo, hi, config=None):
        self.edge_calls.append((lo, hi))
        return [int(self.per_edge[lo:hi].sum())], {"seconds": 1e-3 + 1e-6 * (hi - lo), "split_tasks": 0}

    def count_shard(self, lo, hi, config=None):
        return self.count_shard_edges(int(self.off[lo]), int(self.off[hi]), config)


def test_edge_shards_partition_a_heavy_root():
    """Array-store plans are walked in directed-edge shards: the shards
    partition the requested roots' edges (one heavy root's edges fall into
    several launches) and the raw sums add up."""
    import numpy as np

    from graphautotune_amd.sharded import raw_sharded
    degrees = np.array([200000, 5, 0, 7, 3] + [1] * 100)
    plan = _EdgeStubPlan(degrees)
    raw, shards = raw_sharded(plan, target_s=0.02, first=1)
    assert raw == [int(plan.per_edge.sum())]
    assert [s[0] for s in shards[1:]] == [s[1] for s in shards[:-1]]
    assert shards[0][0] == 0 and shards[-1][1] == int(plan.off[-1])
    assert sum(1 for s in shards if s[1] <= 200000) > 3  # root 0 alone spans several launches
    raw2, shards2 = raw_sharded(plan, target_s=0.02, first=1, roots=(3, 50))
    assert raw2 == [int(plan.per_edge[plan.off[3]:plan.off[50]].sum())]


def test_head_ladder_and_unit_range_estimate():
    """head_ladder times the first edge of roots 65536, ..., 1, 0 and stops
    at the first that reaches the limit; an estimate over a unit range sums
    those units only (exactly, when every block is sampled)."""
    import numpy as np

    from graphautotune_amd.sharded import head_ladder, stratified_estimate

    class Timed(_EdgeStubPlan):
        def count_shard_edges(self, lo, hi, config=None):
            got, st = super().count_shard_edges(lo, hi, config)
            # an edge of root r costs 100 / (1 + r) seconds
            roots = np.searchsorted(self.off, np.arange(lo, hi), side="right") - 1
            st["seconds"] = float((100.0 / (1 + roots)).sum())
            return got, st

    degrees = np.array([3] * 70000)
    plan = Timed(degrees)
    probes, head = head_ladder(plan, limit_s=10.0)
    assert [p["root"] for p in probes] == [65536, 32768, 16384, 8192, 4096, 2048, 1024, 512, 256, 128,
                                           64, 32, 16, 8]
    assert all(p["unit"] == 3 * p["root"] for p in probes)
    assert probes[-1]["seconds"] >= 10.0 and head == 16
    lo = int(plan.off[head])
    est = stratified_estimate(plan, per_stratum=10 ** 9, min_block_s=0.0, units=(lo, int(plan.off[-1])))
    assert est["raw"][0] == pytest.approx(float(plan.per_edge[lo:].sum())) and est["raw_se"][0] == 0.0
    assert est["strata"][0]["units"][0] == lo and est["strata"][-1]["units"][1] == int(plan.off[-1])
    sampled = [c for c in plan.edge_calls[len(probes) + 1:] if c[1] > c[0]]  # not the empty edge_shards checks
    assert sampled and min(c[0] for c in sampled) >= lo


def test_combine_partition_of_root_ranges(tmp_path):
    """--combine adds the raw sums of --roots outputs that partition the
    roots (in any order), divides once and multiplies; gaps, overlaps,
    mismatched workloads and raw sums off the divisor are refused."""
    import json

    from graphautotune_amd.sharded import combine, main
    info = [{"divisor": 2, "multiplier": 3, "relabeled": True, "restrictions": []},
            {"divisor": 1, "multiplier": 1, "relabeled": True, "restrictions": []}]
    base = {"graph": "g", "vertices": 10, "patterns": ["p", "q"], "mode": "unique", "count_info": info}
    a = dict(base, roots=[4, 10], raw=[6, 5], device_seconds=1.5)
    b = dict(base, roots=[0, 4], raw=[4, 7], device_seconds=0.5)
    got = combine([a, b])
    assert got["counts"] == [15, 12] and got["raw"] == [10, 12] and got["device_seconds"] == 2.0
    assert got["roots"] == [[0, 4], [4, 10]]
    with pytest.raises(ValueError):
        combine([a])  # [4, 10) alone leaves a gap
    with pytest.raises(ValueError):
        combine([a, dict(b, roots=[0, 5])])  # overlap
    with pytest.raises(ValueError):
        combine([a, dict(b, patterns=["p"])])
    with pytest.raises(ValueError):
        combine([a, dict(b, raw=[5, 7])])  # 11 is odd
    pa, pb = tmp_path / "a.json", tmp_path / "b.json"
    pa.write_text("log line\n" + json.dumps(a) + "\n")
    pb.write_text(json.dumps(b))
    import contextlib
    import io
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        main(["--combine", str(pa), str(pb)])
    assert json.loads(out.getvalue())["counts"] == [15, 12]
