"""Writes tests/golden/clique_counts.json: whole-graph K3, K4 and K5 counts
of the seeded R-MAT graphs of the clique benchmark, computed by the oracle's
multi-threaded clique counter (oracle/clique_mt.c) on the host, plus a
fingerprint of each graph so a change of the generator is caught.

    python tests/golden/make_clique_counts.py [scale ...]   (default: 20)
"""
import json
import os
import sys
import time
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import graphautotune_amd as gat  # noqa: E402
from oracle import api as orc  # noqa: E402

OUT = os.path.join(HERE, "clique_counts.json")


def fingerprint(off, adj):
    return {"adler32_offsets": zlib.adler32(off.tobytes()), "adler32_adj": zlib.adler32(adj.tobytes())}


def main(scales):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for s in scales:
        g = gat.Graph.rmat(s, 16, 0.57, 0.19, 0.19, 20)
        off, adj = g.csr()
        ro, ra = orc.relabel(off, adj)
        t0 = time.perf_counter()
        k3, k4, k5 = orc.clique_counts_mt(ro, ra, 5, max(1, min(16, os.cpu_count() or 1)))
        dt = time.perf_counter() - t0
        data["rmat%d" % s] = {
            "graph": "R-MAT scale %d, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05), seed 20" % s,
            "vertices": g.v_cnt(), "undirected_edges": g.undirected_edges(), **fingerprint(off, adj),
            "K3": k3, "K4": k4, "K5": k5,
        }
        print("rmat%d: K3 %d K4 %d K5 %d (%.1f s)" % (s, k3, k4, k5, dt), flush=True)
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(x) for x in sys.argv[1:]] or [20])
