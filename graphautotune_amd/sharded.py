"""Count a whole graph as a sequence of root shards (gat_count_shard).

Each shard counts the embeddings whose Schedule vertex 0 maps into a range
of roots; the raw sums of a partition of [0, v_cnt) add up to divisor *
count for every pattern (gat_plan_count_info), so the total is exact
whatever the shard boundaries. Shards grow or shrink so that one launch
takes about `target_s` seconds, which keeps every launch short on skewed
graphs (a few hub roots can carry more work than a million light ones) and
reports progress after each shard. The same loop, over a dealt-out subset
of the ranges, is one GPU's share when a count is split across GPUs.

    python -m graphautotune_amd.sharded --graph cl4194304 --pattern house
"""
import argparse
import json
import sys
import time

from . import core

PATTERNS = {"house": "0100110110010110110010100", "c5": "0100110100010100010110010",
            "k4": "0111101111011110", "k5": "0111110111110111110111110"}


def raw_sharded(plan, config=None, target_s=5.0, first=1, roots=None, log=None):
    """Raw sums of every pattern of `plan` over the roots [lo, hi) (default:
    all of them) and a list of (lo, hi, seconds) shards."""
    lo, n = roots if roots else (0, plan.graph.v_cnt())
    np_ = len(plan.patterns)
    raw = [0] * np_
    shards = []
    size = max(1, first)
    while lo < n:
        hi = min(n, lo + size)
        got, st = plan.count_shard(lo, hi, config)
        raw = [a + b for a, b in zip(raw, got)]
        shards.append((lo, hi, st["seconds"]))
        if log:
            log("shard [%d, %d): %.3f s, split tasks %d; %.1f %% of roots done"
                % (lo, hi, st["seconds"], st["split_tasks"], 100.0 * hi / n))
        t = max(st["seconds"], 1e-4)
        size = max(1, int(size * min(4.0, max(0.25, target_s / t))))
        lo = hi
    return raw, shards


def finish_counts(plan, raw):
    """gat_count's report from raw sums over all roots: raw / divisor *
    multiplier per pattern."""
    infos = [plan.count_info(i) for i in range(len(plan.patterns))]
    for r, info in zip(raw, infos):
        if r % info["divisor"]:
            raise RuntimeError("raw sum %d is not a multiple of the divisor %d" % (r, info["divisor"]))
    return [r // info["divisor"] * info["multiplier"] for r, info in zip(raw, infos)]


def count_sharded(plan, config=None, target_s=5.0, first=1, log=None):
    """Exact counts of every pattern of `plan` (reported as gat_count would)
    and a list of (lo, hi, seconds) shards."""
    raw, shards = raw_sharded(plan, config, target_s, first, None, log)
    return finish_counts(plan, raw), shards


def make_graph(spec):
    """rmatS: R-MAT scale S, edge factor 16, (.57,.19,.19,.05), seed 20;
    clN: Chung-Lu, N vertices, mean degree 24, exponent 2.1, cap 50,000,
    seed 21; or a path to a .bin graph."""
    if spec.startswith("rmat") and spec[4:].isdigit():
        return core.GlobalMemoryGraph.rmat(int(spec[4:]), 16, 0.57, 0.19, 0.19, 20)
    if spec.startswith("cl") and spec[2:].isdigit():
        return core.GlobalMemoryGraph.chung_lu(int(spec[2:]), 24.0, 2.1, 50000, 21)
    return core.GlobalMemoryGraph.load(spec)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--graph", required=True, help="rmatS, clN or a .bin file")
    ap.add_argument("--pattern", action="append", required=True,
                    help="adjacency string or one of %s (repeatable)" % ", ".join(sorted(PATTERNS)))
    ap.add_argument("--mode", default="unique", choices=["unique", "labeled", "reference"])
    ap.add_argument("--config", default=None, help="JSON kernel config")
    ap.add_argument("--target", type=float, default=5.0, help="seconds per shard")
    ap.add_argument("--roots", default=None, metavar="LO:HI",
                    help="count only roots [LO, HI) and report raw sums (add the raw sums of a partition, "
                         "then divide by the divisor and multiply by the multiplier)")
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    g = make_graph(args.graph)
    log("graph %s: %d vertices, %d edges (%.1f s)" % (args.graph, g.v_cnt(), g.undirected_edges(),
                                                   time.perf_counter() - t0))
    pats = [PATTERNS.get(p, p) for p in args.pattern]
    plan = core.Plan(g, pats, args.mode)
    cfg = json.loads(args.config) if args.config else {}
    roots = tuple(int(x) for x in args.roots.split(":")) if args.roots else None
    raw, shards = raw_sharded(plan, cfg, args.target, 1, roots, log)
    dev = sum(s for _, _, s in shards)
    out = {"graph": args.graph, "vertices": g.v_cnt(), "undirected_edges": g.undirected_edges(),
           "patterns": pats, "mode": args.mode, "shards": len(shards), "device_seconds": dev}
    if roots:
        out.update(roots=list(roots), raw=raw, count_info=[plan.count_info(i) for i in range(len(pats))])
    else:
        out.update(counts=finish_counts(plan, raw), edges_per_s=g.undirected_edges() / dev if dev else None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
