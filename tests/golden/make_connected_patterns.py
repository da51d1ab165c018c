"""Writes connected_patterns.txt: every connected graph on 3..6 vertices up to
isomorphism, as the reference's n*n adjacency string (canonical form =
lexicographically largest upper-triangle bit string over all relabellings).

Counts are fixed by enumeration theory (OEIS A001349: 2, 6, 21, 112).
"""
import itertools
import os

import numpy as np


def canon_all(n):
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    m = len(pairs)
    idx = {p: k for k, p in enumerate(pairs)}
    codes = np.arange(1 << m, dtype=np.int64)
    best = codes.copy()
    for perm in itertools.permutations(range(n)):
        # bit k of the relabelled graph comes from bit src[k] of the original
        new = np.zeros_like(codes)
        for k, (i, j) in enumerate(pairs):
            a, b = perm[i], perm[j]
            src = idx[(min(a, b), max(a, b))]
            new |= ((codes >> src) & 1) << k
        best = np.maximum(best, new)
    return pairs, np.unique(best)


def connected(code, n, pairs):
    m = len(pairs)
    adj = [0] * n
    for k, (i, j) in enumerate(pairs):
        if code >> k & 1:
            adj[i] |= 1 << j
            adj[j] |= 1 << i
    seen, frontier = 1, 1
    while frontier:
        nxt = 0
        for v in range(n):
            if frontier >> v & 1:
                nxt |= adj[v]
        frontier = nxt & ~seen
        seen |= nxt
    return seen == (1 << n) - 1, adj


def main():
    out = []
    for n in range(3, 7):
        pairs, cands = canon_all(n)
        for code in sorted(int(c) for c in cands):
            ok, adj = connected(code, n, pairs)
            if ok:
                s = "".join("1" if adj[i] >> j & 1 else "0" for i in range(n) for j in range(n))
                out.append("%d %s" % (n, s))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "connected_patterns.txt")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print(len(out), "patterns ->", path)


if __name__ == "__main__":
    main()
