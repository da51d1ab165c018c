"""The benchmark workloads: seeded graphs and pattern sets of the five
configurations (ER triangles, R-MAT 4/5-cliques, Chung-Lu IEP patterns, the
batched 5-/6-vertex set on R-MAT scale 22, the learned-tuning R-MAT series).

Graph names: "erN" G(N, 16/(N-1)) with seed 1; "rmatS" R-MAT scale S, edge
factor 16, (a,b,c,d) = (.57,.19,.19,.05), seed S (so rmat20 is the clique
benchmark's graph and rmat22 the batched one); "clN" Chung-Lu with N
vertices, mean degree 24, exponent 2.1, degree cap 50,000, seed 21.
"""
import functools
import itertools

import numpy as np

from . import core

TRIANGLE = "011101110"
K4 = "0111101111011110"
K5 = "0111110111110111110111110"
HOUSE = "0100110110010110110010100"
C5 = "0100110100010100010110010"
NAMED = {"triangle": TRIANGLE, "k4": K4, "k5": K5, "house": HOUSE, "c5": C5}


def graph(name):
    if name.startswith("er") and name[2:].isdigit():
        n = int(name[2:])
        return core.GlobalMemoryGraph.erdos_renyi(n, 16.0 / (n - 1), 1)
    if name.startswith("rmat") and name[4:].isdigit():
        s = int(name[4:])
        return core.GlobalMemoryGraph.rmat(s, 16, 0.57, 0.19, 0.19, s)
    if name.startswith("cl") and name[2:].isdigit():
        return core.GlobalMemoryGraph.chung_lu(int(name[2:]), 24.0, 2.1, 50000, 21)
    raise ValueError("unknown graph %r (erN, rmatS or clN)" % name)


def _pairs(n):
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


@functools.lru_cache(maxsize=None)
def connected_patterns(n):
    """Every connected graph on n (<= 6) vertices up to isomorphism, as
    adjacency strings, ordered by canonical code (the canonical code of a
    graph is the largest upper-triangle bit code over all relabellings)."""
    pairs = _pairs(n)
    idx = {p: k for k, p in enumerate(pairs)}
    codes = np.arange(1 << len(pairs), dtype=np.int64)
    best = codes.copy()
    for perm in itertools.permutations(range(n)):
        new = np.zeros_like(codes)
        for k, (i, j) in enumerate(pairs):
            a, b = perm[i], perm[j]
            new |= ((codes >> idx[(min(a, b), max(a, b))]) & 1) << k
        np.maximum(best, new, out=best)
    out = []
    for code in sorted(int(c) for c in np.unique(best)):
        adj = [0] * n
        for k, (i, j) in enumerate(pairs):
            if code >> k & 1:
                adj[i] |= 1 << j
                adj[j] |= 1 << i
        seen = frontier = 1
        while frontier:
            nxt = 0
            for v in range(n):
                if frontier >> v & 1:
                    nxt |= adj[v]
            frontier = nxt & ~seen
            seen |= nxt
        if seen == (1 << n) - 1:
            out.append("".join("1" if adj[i] >> j & 1 else "0" for i in range(n) for j in range(n)))
    return tuple(out)


def edge_density(p):
    n = int(round(len(p) ** 0.5))
    return sum(p[i * n + j] == "1" for i, j in _pairs(n)) / (n * (n - 1) / 2)


def batched_patterns(seed=6, n_six=10, density=(0.4, 0.8)):
    """The batched configuration's pattern set: all 21 connected 5-vertex
    patterns plus n_six connected 6-vertex patterns of edge density within
    `density`, drawn without replacement by numpy's default_rng(seed)."""
    six = [p for p in connected_patterns(6) if density[0] <= edge_density(p) <= density[1]]
    rng = np.random.default_rng(seed)
    return list(connected_patterns(5)) + [six[i] for i in rng.choice(len(six), n_six, replace=False)]
