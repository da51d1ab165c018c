"""Pins for the CPU oracle itself: every oracle function is checked against
closed forms, brute force or an independent derivation, chosen so that a
plausible mistake (a dropped term, a wrong sign, a transposed index) fails.
"""
import itertools
import math

import numpy as np
import pytest

from patterns import C5, HOUSE, K4, K5, P3, TRIANGLE, clique, connected_patterns


def csr_of_edges(n, edges):
    adj = [[] for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].append(v)
            adj[v].append(u)
    off = [0]
    flat = []
    for r in adj:
        r = sorted(set(r))
        flat += r
        off.append(len(flat))
    return np.array(off, dtype=np.int64), np.array(flat, dtype=np.int32)


def complete(m):
    return csr_of_edges(m, [(i, j) for i in range(m) for j in range(i + 1, m)])


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    return csr_of_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p])


def to_dense(off, adj):
    n = len(off) - 1
    A = np.zeros((n, n), dtype=np.int64)
    for v in range(n):
        A[v, adj[off[v]:off[v + 1]]] = 1
    return A


def brute_labeled(off, adj, pattern):
    """All injective maps, by definition (tiny graphs only)."""
    n = len(off) - 1
    A = to_dense(off, adj)
    k = int(round(len(pattern) ** 0.5))
    pe = [(i, j) for i in range(k) for j in range(i + 1, k) if pattern[i * k + j] == "1" or pattern[j * k + i] == "1"]
    return sum(all(A[m[i], m[j]] for i, j in pe) for m in itertools.permutations(range(n), k))


# ---- schedule.hpp:149-170 (connected labelled graphs by edge parity) ----
def test_graph_cnt_closed_forms(orc):
    connected_labelled = [1, 1, 4, 38, 728]  # OEIS A001187, k = 1..5
    for k in range(1, 6):
        e, o = orc.graph_cnt(k)
        assert e + o == connected_labelled[k - 1]
        assert e - o == (-1) ** (k - 1) * math.factorial(k - 1)  # Moebius of the partition lattice


# ---- schedule.hpp:173-191 ----
def test_automorphisms_by_orbit_counting(orc):
    """|Aut| = n! / #distinct relabellings (orbit-stabiliser), an independent
    derivation from the set of images rather than the stabiliser."""
    for n in (3, 4, 5):
        for p in connected_patterns(n):
            images = set()
            for perm in itertools.permutations(range(n)):
                s = ["0"] * (n * n)
                for i in range(n):
                    for j in range(n):
                        if p[i * n + j] == "1":
                            s[perm[i] * n + perm[j]] = "1"
                images.add("".join(s))
            assert orc.automorphisms(p) * len(images) == math.factorial(n)
    assert orc.automorphisms(C5) == 10
    assert orc.automorphisms(HOUSE) == 2
    assert orc.automorphisms(clique(5)) == 120


# ---- labelled embeddings: closed forms and brute force ----
def test_labeled_closed_forms(orc):
    for m in (5, 7, 9):
        off, adj = complete(m)
        for p in (TRIANGLE, K4, K5, HOUSE, C5):
            k = int(round(len(p) ** 0.5))
            assert orc.labeled_count(off, adj, p) == math.perm(m, k)
        assert orc.unique_count(off, adj, C5) == 12 * math.comb(m, 5)
        assert orc.unique_count(off, adj, HOUSE) == 60 * math.comb(m, 5)
    # complete bipartite: no odd cycle, C4 = C(a,2) C(b,2)
    a, b = 4, 5
    off, adj = csr_of_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])
    for p in (TRIANGLE, K4, C5, HOUSE):
        assert orc.labeled_count(off, adj, p) == 0
    c4 = "0101101001011010"
    assert orc.unique_count(off, adj, c4) == math.comb(a, 2) * math.comb(b, 2)


def test_triangles_trace(orc):
    off, adj = random_graph(40, 0.2, 3)
    A = to_dense(off, adj)
    assert orc.unique_count(off, adj, TRIANGLE) == np.trace(A @ A @ A) // 6


def test_labeled_equals_brute_force_and_naive(orc):
    for seed in range(3):
        off, adj = random_graph(8, 0.5, seed)
        for p in (TRIANGLE, K4, HOUSE, C5, "0110101111010110"):
            want = brute_labeled(off, adj, p)
            assert orc.labeled_count(off, adj, p) == want
            assert orc.naive_count(off, adj, p) == want  # naive_pm.cu restatement


# ---- reference engine semantics ----
def test_reference_cliques_are_k_times_unique(orc):
    from graphautotune_amd._lib import ScheduleData
    off, adj = random_graph(30, 0.4, 5)
    for k in (3, 4, 5):
        u = orc.unique_count(off, adj, clique(k))
        r, ok = orc.reference_count(off, adj, orc.schedule(clique(k), ScheduleData))
        assert ok and r == k * u


def _reference_brute(off, adj, s):
    """The reference engine's value written from its definition by a separate
    enumeration over all tuples (v_0..v_{b-1}) (tiny graphs only):
      * dup(l): v_l equals an earlier vertex; every prefix set computed at a
        dup level is emptied (engine_kernel.cuh:243-258), so a prefix set is
        empty when any level of its chain is a dup, else the intersection of
        the neighbourhoods of its vertices;
      * v_l must lie in its loop set and below v_f for each filtered
        restriction (f, l) (engine_kernel.cuh:290-338);
      * IEP factors: size-only prefixes exclude the vertices chosen before
        their last level (array_vertex_set.cuh:248-275), a size-only
        one-vertex prefix is the plain degree, other prefixes subtract every
        chosen vertex found in them (array_vertex_set.cuh:318-331)."""
    n = len(off) - 1
    N = [set(adj[off[v]:off[v + 1]].tolist()) for v in range(n)]
    b = s.basic_vertexes
    pref = [list(s.prefix_data[i][:s.prefix_depth[i]]) for i in range(s.total_prefix_num)]
    so = list(s.prefixs_size_only[:s.total_prefix_num])
    restr = list(zip(s.restrictions[:s.total_restrictions_num], s.restriction_second[:s.total_restrictions_num]))
    total = 0
    for tup in itertools.product(range(n), repeat=b):
        dup = [any(tup[l] == tup[i] for i in range(l)) for l in range(b)]

        def setof(d):
            if any(dup[j] for j in d):
                return set()
            return set.intersection(*[N[tup[j]] for j in d])

        ok = True
        for l in range(1, b):
            if tup[l] not in setof(pref[s.loop_set_prefix_id[l]]):
                ok = False
                break
            if any(t == l and not tup[l] < tup[f] for f, t in restr):
                ok = False
                break
        if not ok:
            continue
        ans = []
        for pid in s.iep_vertex_id[:s.iep_prefix_num]:
            d = pref[pid]
            S = setof(d)
            if not S:
                ans.append(0)
            elif so[pid] and len(d) > 1:
                ans.append(len(S - {tup[i] for i in range(d[-1])}))
            elif so[pid]:
                ans.append(len(S))
            else:
                ans.append(len(S) - sum(tup[i] in S for i in range(b)))
        val, last, local = 1, -1, 0
        for g in range(s.subgroups_num):
            a = ans[s.iep_ans_pos[g]]
            val = a if g == last + 1 else val * a
            if s.iep_flag[g]:
                last = g
                local += val * s.iep_coef[g]
        total += local
    return total % (1 << 64)


def test_reference_matches_definition_brute_force(orc):
    from graphautotune_amd._lib import ScheduleData
    for seed in range(2):
        off, adj = random_graph(9, 0.5, 10 + seed)
        for p in (TRIANGLE, K4, HOUSE, C5, P3):
            s = orc.schedule(p, ScheduleData)
            got, ok = orc.reference_count(off, adj, s)
            assert ok
            assert got == _reference_brute(off, adj, s), p


def test_house_reference_equals_unique(orc):
    """For the house the emptied duplicate branches contribute 0 and all IEP
    prefixes end at the last basic vertex, so reference == unique."""
    from graphautotune_amd._lib import ScheduleData
    off, adj = random_graph(25, 0.35, 4)
    r, ok = orc.reference_count(off, adj, orc.schedule(HOUSE, ScheduleData))
    assert ok and r == orc.unique_count(off, adj, HOUSE)


def test_golden_intersection_15(orc):
    # tests/test_vertex_set.cu:44-48,76 as a graph (see test_gpu_parity)
    a, b = 61, 62
    edges = [(a, i) for i in range(1, 31)] + [(b, 2 * i) for i in range(1, 31)] + [(a, b)]
    off, adj = csr_of_edges(63, edges)
    assert orc.unique_count(off, adj, TRIANGLE) == 15


def test_relabel_and_clique_walg(orc):
    off, adj = random_graph(60, 0.15, 2)
    ro, ra = orc.relabel(off, adj)
    deg = np.diff(ro)
    assert np.all(deg[:-1] >= deg[1:])  # descending degree order
    assert sorted(np.diff(off).tolist()) == sorted(deg.tolist())
    for k in (3, 4, 5):
        c, w = orc.clique_walg(ro, ra, k)
        assert c == orc.unique_count(off, adj, clique(k))
        # W_alg by its definition, recomputed independently in Python
        N = [ra[ro[v]:ro[v + 1]] for v in range(len(ro) - 1)]
        S = lambda v: [x for x in N[v] if x < v]  # noqa: E731
        tot = 0

        def rec(Sprev, d, chosen):
            nonlocal tot
            for v in Sprev:
                A = S(v)
                B = [x for x in Sprev if x < v]
                out = sorted(set(A) & set(B))
                if d == k - 2:
                    tot += 4 * (len(A) + len(B)) + 16
                else:
                    tot += 4 * (len(A) + len(B)) + 4 * len(out) + 16
                    rec(out, d + 1, chosen + [v])

        for v0 in range(len(ro) - 1):
            for v1 in S(v0):
                A, B = S(v1), [x for x in S(v0) if x < v1]
                out = sorted(set(A) & set(B))
                if k == 3:
                    tot += 4 * (len(A) + len(B)) + 32
                else:
                    tot += 4 * (len(A) + len(B)) + 4 * len(out) + 32
                    rec(out, 2, [v0, v1])
        assert w == tot


def test_clique_bitmap_walg_definition(orc):
    """Counts of the bitmap-store restatement equal the brute-force unique
    counts, and its W_alg equals the definition recomputed here: per v0 with
    d = |N+(v0)| >= 2, 4d + 16; per position p, 4(|N+(L[p])| + p) + 16 +
    4 w(p); for each pair, triple and quadruple of mutually adjacent
    positions that the deeper levels visit, 8 w(smallest position), with
    w(x) = x // 32 + 1."""
    off, adj = random_graph(45, 0.3, 7)
    ro, ra = orc.relabel(off, adj)
    n = len(ro) - 1
    up = [[x for x in ra[ro[v]:ro[v + 1]] if x < v] for v in range(n)]
    for k in (3, 4, 5, 6):
        c, w = orc.clique_bitmap_walg(ro, ra, k)
        assert c == orc.unique_count(off, adj, clique(k))
        tot = 0
        for v0 in range(n):
            L = up[v0]
            d = len(L)
            if d < 2:
                continue
            tot += 4 * d + 16 + sum(4 * (len(up[x]) + p) + 16 + 4 * (p // 32 + 1) for p, x in enumerate(L))
            adjL = {(p, q) for p in range(d) for q in range(p) if L[q] in set(up[L[p]])}
            # cliques of size j among positions, visited at depth j (j = 2, 3, 4)
            for size in (2, 3, 4):
                if k < size + 2:
                    break
                for combo in itertools.combinations(range(d), size):
                    if all((b, a) in adjL for a, b in itertools.combinations(combo, 2)):
                        tot += 8 * (min(combo) // 32 + 1)
        assert w == tot, (k, w, tot)


def test_clique_bitmap_walg_sample_partitions(orc):
    off, adj = random_graph(80, 0.2, 11)
    ro, ra = orc.relabel(off, adj)
    for k in (3, 5):
        full = orc.clique_bitmap_walg(ro, ra, k)
        parts = [orc.clique_bitmap_walg_sample(ro, ra, k, 4, f) for f in range(4)]
        assert (sum(p[0] for p in parts), sum(p[1] for p in parts)) == full


def brute_restricted(off, adj, rows, restrictions, lo, hi):
    """Every injective map of the pattern's vertices, filtered (pure Python)."""
    n, k = len(off) - 1, len(rows)
    nb = [set(adj[off[v]:off[v + 1]].tolist()) for v in range(n)]
    total = 0
    for psi in itertools.permutations(range(n), k):
        if not lo <= psi[0] < hi:
            continue
        if any((rows[i] >> j) & 1 and psi[j] not in nb[psi[i]] for i in range(k) for j in range(i)):
            continue
        if all(psi[b] < psi[a] for a, b in restrictions):
            total += 1
    return total


def rows_of(adj_str):
    n = int(round(len(adj_str) ** 0.5))
    return [sum(1 << j for j in range(n) if adj_str[i * n + j] == "1") for i in range(n)]


def test_restricted_count_direction_by_hand(orc):
    # K3 with psi(1) < psi(0) and psi(2) < psi(1): only psi = (2, 1, 0), rooted at 2
    off, adj = complete(3)
    rows, R = rows_of(TRIANGLE), [(0, 1), (1, 2)]
    assert [orc.restricted_count(off, adj, rows, R, v, v + 1) for v in range(3)] == [0, 0, 1]
    assert orc.restricted_count(off, adj, rows, []) == 6  # |Aut(K3)| labeled maps


@pytest.mark.parametrize("seed", [1, 2])
def test_restricted_count_against_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    n = 8
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.55]
    off, adj = csr_of_edges(n, edges)
    cases = [
        (TRIANGLE, []), (TRIANGLE, [(0, 1), (1, 2)]), (TRIANGLE, [(0, 1), (0, 2)]),
        (P3, [(0, 2)]),
        ("001001110", [(0, 1)]),  # path 0-2-1: vertex 1 has no earlier neighbour
        (K4, [(0, 1), (1, 2), (2, 3)]), (K4, [(2, 3)]),
        (HOUSE, [(1, 2)]), (HOUSE, []),
        (C5, [(2, 3), (0, 3), (0, 1), (0, 2), (0, 4)]), (C5, [(3, 0), (4, 1)]),
    ]
    for pat, R in cases:
        rows = rows_of(pat)
        for lo, hi in ((0, n), (0, 3), (3, 8), (2, 5), (7, 8), (4, 4)):
            want = brute_restricted(off, adj, rows, R, lo, hi)
            assert orc.restricted_count(off, adj, rows, R, lo, hi) == want, (pat, R, lo, hi)


def test_reference_count_roots_partitions_the_count(orc):
    """Every first vertex starts its own walk, so the root-range form adds up
    over a partition to the full reference count (pinned itself by k * unique
    for cliques and by the brute-force restatement above)."""
    from graphautotune_amd._lib import ScheduleData
    rng = np.random.default_rng(4)
    n = 40
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.2]
    off, adj = csr_of_edges(n, edges)
    for p in (TRIANGLE, K4, HOUSE, C5, P3):
        s = orc.schedule(p, ScheduleData)
        full, ok = orc.reference_count(off, adj, s)
        if not ok:
            continue
        cuts = [0, 1, 7, 20, 39, 40]
        parts = [orc.reference_count_roots(off, adj, s, lo, hi) for lo, hi in zip(cuts, cuts[1:])]
        assert all(o for _, o in parts) and sum(v for v, _ in parts) == full, p


def test_clique_counts_mt_against_brute_force_and_closed_forms(orc):
    """The whole-graph clique counter (all host cores): brute-force unique
    counts on random graphs in input and degree order (the count does not
    depend on the labelling), C(m, k) on complete graphs (rows wider than
    one 64-bit word at m = 150), nothing on a bipartite graph, and stride
    samples that partition the roots; any thread count gives one result."""
    for seed, (n, p) in enumerate([(40, 0.3), (70, 0.2), (90, 0.12)]):
        off, adj = random_graph(n, p, 20 + seed)
        want = [orc.unique_count(off, adj, clique(k)) for k in (3, 4, 5, 6)]
        assert orc.clique_counts_mt(off, adj, 6, 3) == want
        ro, ra = orc.relabel(off, adj)
        assert orc.clique_counts_mt(ro, ra, 6, 1) == want
        assert orc.clique_counts_mt(ro, ra, 4, 8) == want[:2]
        parts = [orc.clique_counts_mt(ro, ra, 6, 2, first=f, stride=5) for f in range(5)]
        assert [sum(c) for c in zip(*parts)] == want
    for m in (7, 66, 150):
        off, adj = complete(m)
        assert orc.clique_counts_mt(off, adj, 6, 4) == [math.comb(m, k) for k in (3, 4, 5, 6)]
    a, b = 9, 13
    off, adj = csr_of_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])
    assert orc.clique_counts_mt(off, adj, 6, 2) == [0, 0, 0, 0]
    off, adj = csr_of_edges(5, [])
    assert orc.clique_counts_mt(off, adj, 5, 2) == [0, 0, 0]


def test_golden_clique_counts_small_scales(gat, orc):
    """The committed clique counts of R-MAT scales 12 and 16 are what the
    oracle computes now (scale 20 is checked against the GPU and by the
    benchmark's CPU baseline)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "clique_counts.json")) as f:
        gold = json.load(f)
    for scale in (12, 16):
        g = gat.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, 20)  # csr() views g's memory: keep g alive
        off, adj = g.csr()
        want = gold["rmat%d" % scale]
        assert orc.clique_counts_mt(off, adj, 5, 4) == [want["K3"], want["K4"], want["K5"]]
    g = gat.Graph.rmat(12, 16, 0.57, 0.19, 0.19, 20)
    off, adj = g.csr()
    assert orc.unique_count(off, adj, K4) == gold["rmat12"]["K4"]
