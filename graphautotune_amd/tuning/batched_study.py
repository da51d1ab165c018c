"""Batched plan of all 21 connected 5-vertex patterns plus 10 seeded random
connected 6-vertex patterns on an R-MAT graph: per-pattern Schedule ranks
tuned on a smaller R-MAT proxy (schedules.tune_schedule_ranks), then the
batched tuned plan against the plain single-pattern plans with the
reference's Schedules, at identical counts.

    python -m graphautotune_amd.tuning.batched_study --proxy-scale 10 --scale 12

The pattern set is workloads.batched_patterns(): the 21 connected 5-vertex
patterns and 10 of the connected 6-vertex ones of edge density 0.4-0.8,
drawn by numpy's default_rng(6).
"""
import argparse
import json
import sys
import time

from .. import workloads
from ..core import Graph, Plan
from .schedules import tune_schedule_ranks

def connected_patterns(n):
    return list(workloads.connected_patterns(n))


def pattern_set():
    return workloads.batched_patterns()


def timed(plan, config=None):
    plan.count(config)
    counts, st = plan.count(config)
    return counts, st["seconds"]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--proxy-scale", type=int, default=10)
    ap.add_argument("--scale", type=int, default=12)
    ap.add_argument("--max-rank", type=int, default=12)
    ap.add_argument("--out", default=None, help="write the result as JSON here")
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    pats = pattern_set()
    proxy = Graph.rmat(args.proxy_scale, 16, 0.57, 0.19, 0.19, 22)
    t0 = time.perf_counter()
    ranks, hist = tune_schedule_ranks(proxy, pats, max_rank=args.max_rank, log=log)
    tune_s = time.perf_counter() - t0
    g = Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, 22)
    cb, tb = timed(Plan(g, pats, "unique", schedule="ranked", schedule_rank=ranks))
    cr, tr = timed(Plan(g, pats, "unique"))
    if cb != cr:
        raise AssertionError("tuned schedules changed the counts")
    plain = []
    for i, p in enumerate(pats):
        c, t = timed(Plan(g, [p], "unique"))
        if c[0] != cb[i]:
            raise AssertionError("pattern %d: plain %d vs batched %d" % (i, c[0], cb[i]))
        plain.append(t)
    out = {"patterns": pats, "proxy": "R-MAT s%d" % args.proxy_scale, "graph": "R-MAT s%d" % args.scale,
           "vertices": g.v_cnt(), "undirected_edges": g.undirected_edges(), "ranks": ranks,
           "tuning_seconds": tune_s, "batched_tuned_ms": tb * 1e3, "batched_reference_ms": tr * 1e3,
           "plain_reference_ms": sum(plain) * 1e3, "batched_tuned_vs_plain": sum(plain) / tb, "counts": cb}
    print(json.dumps(out), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
