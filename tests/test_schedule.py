"""Schedule builder: product (libgat, C++) against the oracle's literal
restatement of include/core/schedule.hpp, plus invariants that pin both."""
import itertools
import math

import pytest

from patterns import C5, HOUSE, K4, K5, P3, P4, TRIANGLE, clique, connected_patterns

# The two 6-vertex patterns whose unmemoised restriction search (the
# oracle's literal restatement) takes minutes: C6 and the triangular prism.
SLOW = {"011100101010110001100011010101001110", "011000100100100010010001001001000110"}


def all_patterns():
    out = []
    for n in (3, 4, 5, 6):
        out += [p for p in connected_patterns(n) if p not in SLOW]
    return out


def test_product_equals_oracle_on_all_small_patterns(gat, orc):
    from graphautotune_amd._lib import ScheduleData
    pats = all_patterns()
    assert len(pats) == 2 + 6 + 21 + 112 - len(SLOW)
    for p in pats:
        s = gat.Schedule(p)
        o = orc.schedule(p, ScheduleData)
        assert bytes(s.data) == bytes(o), p


def test_reference_strings(gat, orc):
    from graphautotune_amd._lib import ScheduleData
    for p in (P3, P4):  # scripts/run_pattern_matching.sh:19,21 (6 and 7 vertices)
        assert bytes(gat.Schedule(p).data) == bytes(orc.schedule(p, ScheduleData))


def n_linear_extensions(n, pairs):
    return sum(all(perm[a] < perm[b] for a, b in pairs) for perm in itertools.permutations(range(n)))


def bell(k):
    b = [[1]]
    for i in range(1, k + 1):
        row = [b[-1][-1]]
        for x in b[-1]:
            row.append(row[-1] + x)
        b.append(row)
    return b[k][0]


@pytest.mark.parametrize("p", [TRIANGLE, K4, K5, HOUSE, C5, P3] + connected_patterns(5))
def test_schedule_invariants(gat, p):
    s = gat.Schedule(p)
    n = s.data.v_cnt
    q = [[(s.data.adj_rows[i] >> j) & 1 for j in range(n)] for i in range(n)]
    # the ordered pattern is the input relabelled by perm
    pin = gat.Pattern(p)
    for i in range(n):
        for j in range(n):
            assert q[s.perm[i]][s.perm[j]] == int(pin.has_edge(i, j))
    # restriction set breaks all symmetry on K_n: n!/|Aut| linear extensions
    assert n_linear_extensions(n, s.all_restrictions) == math.factorial(n) // s.automorphisms
    # the IEP suffix is maximal over valid orders, and independent
    def suffix(rows):
        for k in range(2, n - 1):
            v = n - k
            if any(rows[v][w] for w in range(v + 1, n)):
                return k - 1
        return n - 2
    best = 0
    for perm in itertools.permutations(range(n)):
        r = [[0] * n for _ in range(n)]
        for i in range(n):
            for j in range(n):
                r[perm[i]][perm[j]] = int(pin.has_edge(i, j))
        if all(any(r[i][j] for j in range(i)) for i in range(1, n)):
            best = max(best, suffix(r))
    assert s.iep_suffix_vertexes == best
    b = s.basic_vertexes
    assert b == n - s.iep_suffix_vertexes
    # loop set of vertex i = its earlier neighbours
    for i in range(1, b):
        assert s.prefixs[s.loop_set_prefix_id[i]] == [j for j in range(i) if q[i][j]]
    # prefixes sorted by last vertex; father = prefix minus last; size-only = leaf
    lasts = [pr[-1] for pr in s.prefixs]
    assert lasts == sorted(lasts)
    for k, pr in enumerate(s.prefixs):
        f = s.prefixs_father[k]
        assert (f == -1) == (len(pr) == 1)
        if f != -1:
            assert s.prefixs[f] == pr[:-1]
        assert s.prefixs_size_only[k] == int(k not in s.prefixs_father)
    # IEP: Bell(suffix) partitions, coefficient = prod (-1)^(|B|-1) (|B|-1)!
    flags = [i for i, f in enumerate(s.iep_flag) if f]
    assert len(flags) == bell(s.iep_suffix_vertexes)
    coefs = sorted(s.iep_coef[i] for i in flags)
    want = []

    def parts(seq):
        if not seq:
            yield []
            return
        first, rest = seq[0], seq[1:]
        for p_ in parts(rest):
            yield [[first]] + p_
            for i in range(len(p_)):
                yield p_[:i] + [[first] + p_[i]] + p_[i + 1:]
    for part in parts(list(range(s.iep_suffix_vertexes))):
        c = 1
        for blk in part:
            c *= (-1) ** (len(blk) - 1) * math.factorial(len(blk) - 1)
        want.append(c)
    assert coefs == sorted(want)
    # filtered restrictions = those inside the basic vertices, sorted by (second, first)
    assert s.restrictions == sorted([r for r in s.all_restrictions if r[1] < b], key=lambda r: (r[1], r[0]))


def test_cliques_use_identity_and_all_pairs(gat):
    for k in (3, 4, 5, 6):
        s = gat.Schedule(clique(k))
        assert s.perm == list(range(k))
        assert sorted(s.all_restrictions) == sorted((i, j) for i in range(k) for j in range(i + 1, k))
        assert s.iep_suffix_vertexes == 1


def test_bad_patterns_rejected(gat):
    from graphautotune_amd._lib import GatError
    for bad in ("01", "0110x", "0" * 5, "010100000", "0100100000000000"):  # not square / chars / disconnected
        with pytest.raises((GatError, ValueError)):
            gat.Schedule(bad)


# ---- Schedule candidates (graph-aware choice for unique/labeled plans) ----

def _restrictions(s):
    d = s.data
    return sorted((d.all_restr_first[i], d.all_restr_second[i]) for i in range(d.n_all_restrictions))


def test_candidate_zero_is_the_reference_schedule(gat):
    for p in all_patterns():
        cands, total = gat.schedule_candidates(p, limit=1)
        assert total >= 1 and bytes(cands[0].data) == bytes(gat.Schedule(p).data), p


@pytest.mark.parametrize("p", [HOUSE, C5, K4] + connected_patterns(5)[::4] + connected_patterns(6)[::23])
def test_every_candidate_is_a_valid_schedule(gat, p):
    """Each candidate relabels the pattern into a valid order with the largest
    IEP suffix, carries a restriction set that leaves exactly n!/|Aut|
    orderings (the validity test of schedule.hpp:346-357, by brute force),
    and the list is sorted by cost with no two candidates alike."""
    ref = gat.Schedule(p)
    n = ref.data.v_cnt
    aut = gat.Pattern(p).automorphisms()
    for stats in ((5000, 10 ** 6, 10 ** 4), (65536, 1819096, 2855038)):
        cands, total = gat.schedule_candidates(p, *stats, limit=512)
        assert len(cands) == total
        seen = set()
        for k, s in enumerate(cands):
            d = s.data
            assert d.iep_suffix_vertexes == ref.data.iep_suffix_vertexes
            rows = tuple(d.adj_rows[i] for i in range(n))
            for i in range(1, n):  # every vertex after the first has an earlier neighbour
                assert rows[i] & ((1 << i) - 1)
            # the ordered pattern is the input relabelled by perm
            pin = gat.Pattern(p)
            for i in range(n):
                for j in range(n):
                    assert pin.has_edge(i, j) == bool(rows[d.perm[i]] >> d.perm[j] & 1)
            R = _restrictions(s)
            assert n_linear_extensions(n, R) == math.factorial(n) // aut
            key = (rows, tuple(R))
            assert key not in seen
            seen.add(key)
            if k:
                assert cands[k - 1].data.best_cost <= d.best_cost


def test_graph_schedule_policy_rules(gat):
    from graphautotune_amd._lib import GAT_EINVAL, GatError
    g = gat.Graph.erdos_renyi(40, 0.3, 3)
    with pytest.raises(GatError) as e:
        gat.Plan(g, [HOUSE], "reference", schedule="graph")
    assert e.value.code == GAT_EINVAL
    with pytest.raises(GatError):
        gat.Plan(g, [HOUSE], "unique", schedule="graph", schedule_rank=-1)
    # a rank past the end takes the last candidate
    _, total = gat.schedule_candidates(HOUSE, g.v_cnt(), 2 * g.undirected_edges(), g.triangles())
    last = gat.Plan(g, [HOUSE], "unique", schedule="graph", schedule_rank=total - 1).schedule(0)
    past = gat.Plan(g, [HOUSE], "unique", schedule="graph", schedule_rank=total + 5).schedule(0)
    assert bytes(last.data) == bytes(past.data)


def test_per_pattern_schedule_ranks(gat):
    g = gat.Graph.erdos_renyi(50, 0.2, 1)
    stats = (g.v_cnt(), 2 * g.undirected_edges(), g.triangles())
    pats = [HOUSE, C5]
    plan = gat.Plan(g, pats, "unique", schedule="graph", schedule_rank=[3, 7])
    for p, r, i in zip(pats, (3, 7), range(2)):
        cands, total = gat.schedule_candidates(p, *stats, limit=64)
        want = cands[min(r, total - 1)]
        assert list(plan.schedule(i).data.perm) == list(want.data.perm)
    with pytest.raises(ValueError):
        gat.Plan(g, [HOUSE], "unique", schedule="graph", schedule_rank=[1, 2])


def test_ranked_policy_is_graph_independent(gat):
    """A "ranked" rank names the same Schedule on every graph: the candidates
    ranked under the reference's fixed statistics; rank 0 is its pick."""
    cands, total = gat.schedule_candidates(C5, limit=64)
    graphs = [gat.Graph.erdos_renyi(50, 0.2, 1), gat.Graph.rmat(8, 8, 0.57, 0.19, 0.19, 1)]
    for r in (0, 5, total - 1, total + 3):
        for g in graphs:
            s = gat.Plan(g, [C5], "unique", schedule="ranked", schedule_rank=r).schedule(0)
            assert bytes(s.data) == bytes(cands[min(r, total - 1)].data)
    assert bytes(cands[0].data) == bytes(gat.Schedule(C5).data)
