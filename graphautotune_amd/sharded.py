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
    python -m graphautotune_amd.sharded --graph rmat22 --batched-set --estimate
    python -m graphautotune_amd.sharded --combine part_a.json part_b.json
"""
import argparse
import json
import sys
import time

from . import core, workloads

PATTERNS = workloads.NAMED


def edge_shards(plan, config=None):
    """Whether the plan's count can be split by directed edges
    (gat_count_shard_edges: the array store). Clique plans on the bitmap
    store are split by top vertex."""
    from ._lib import GAT_EUNSUPPORTED, GatError
    try:
        plan.count_shard_edges(0, 0, config)
        return True
    except GatError as e:
        if e.code == GAT_EUNSUPPORTED:
            return False
        raise


def raw_sharded(plan, config=None, target_s=5.0, first=1, roots=None, log=None, quiet_below=0.0):
    """Raw sums of every pattern of `plan` over the roots [lo, hi) (default:
    all of them) and a list of (lo, hi, seconds) shards. Launches are sized
    to about target_s each: by directed edges of the CSR the plan runs on for
    the array store (so one heavy root's work spreads over launches), by
    roots for the bitmap clique store; the first shard has `first` units and
    each next one grows or shrinks toward target_s (at most 4x). Shard
    bounds are edge indices or root ids accordingly. Shards shorter than
    quiet_below seconds are not logged."""
    r_lo, r_hi = roots if roots else (0, plan.graph.v_cnt())
    if edge_shards(plan, config):
        off = plan.run_offsets()
        lo, n = int(off[r_lo]), int(off[r_hi])
        size, unit = max(1, first), "edges"
        count = lambda a, b: plan.count_shard_edges(a, b, config)  # noqa: E731
    else:
        lo, n = r_lo, r_hi
        size, unit = max(1, first), "roots"
        count = lambda a, b: plan.count_shard(a, b, config)  # noqa: E731
    start = lo
    np_ = len(plan.patterns)
    raw = [0] * np_
    shards = []
    while lo < n:
        hi = min(n, lo + size)
        got, st = count(lo, hi)
        raw = [a + b for a, b in zip(raw, got)]
        shards.append((lo, hi, st["seconds"]))
        if log and st["seconds"] >= quiet_below:
            log("shard [%d, %d) %s: %.3f s, split tasks %d; %.1f %% done; raw sums so far %s"
                % (lo, hi, unit, st["seconds"], st["split_tasks"], 100.0 * (hi - start) / max(1, n - start), raw))
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


def strata_of(n, first=1):
    """Root strata [lo, hi) of a degree-ordered graph: [0, first), then
    ranges that double in size, so each holds roots of similar degree."""
    cuts, lo, size = [], 0, max(1, first)
    while lo < n:
        hi = min(n, lo + size)
        cuts.append((lo, hi))
        lo, size = hi, size * 2
    return cuts


def unit_space(plan, config=None):
    """(units, unit name, count function, first unit of each root) of the
    plan's sampling units: directed edges of the CSR the plan runs on for
    the array store, roots for the bitmap clique store."""
    if edge_shards(plan, config):
        off = plan.run_offsets()
        return (int(off[-1]), "edges", lambda a, b: plan.count_shard_edges(a, b, config),
                lambda r: int(off[r]))
    return plan.graph.v_cnt(), "roots", lambda a, b: plan.count_shard(a, b, config), int


def head_ladder(plan, config=None, limit_s=20.0, start_root=65536, log=None):
    """Times single units at the heavy end of the degree order: the first
    unit of roots start_root, start_root/2, ..., 1, 0 (for the array store
    the edge from that root to its highest-degree neighbour, the root's most
    expensive edge). Stops after the first unit that takes limit_s or more,
    or at root 0. A unit is one launch and cannot be split further, so a
    launch of the units before the last root under the limit may run for
    minutes. Returns ([{"root", "unit", "seconds"}...], the smallest probed
    root whose unit stayed under limit_s; 0 when all did)."""
    n, unit, count, first = unit_space(plan, config)
    r = min(start_root, plan.graph.v_cnt() - 1)
    probes, ok = [], r + 1
    while True:
        u = first(r)
        if u < n:
            _, st = count(u, u + 1)
            probes.append({"root": r, "unit": u, "seconds": st["seconds"]})
            if log:
                log("head probe: root %d, %s [%d, %d): %.3f s" % (r, unit, u, u + 1, st["seconds"]))
            if st["seconds"] >= limit_s:
                break
            ok = r
        if r == 0:
            break
        r //= 2
    return probes, ok


def stratified_estimate(plan, config=None, per_stratum=8, min_block_s=1.0, seed=0, log=None, units=None):
    """Estimated whole-graph raw sums and device time of `plan` from a
    stratified sample of blocks, with standard errors.

    The sampling units are the directed edges of the CSR the plan runs on
    for the array store (gat_count_shard_edges) and the roots for the bitmap
    clique store (gat_count_shard). Both are in degree order (unique and
    labeled plans run on the degree-ordered graph; an edge's index follows
    its root's degree, then its neighbour's), so the strata of strata_of()
    over the unit indices hold units of similar cost; units of a hub root
    fall into strata of their own. Each stratum is cut into equal blocks of
    consecutive units, the block size grown from one unit until a block,
    drawn at random, takes at least min_block_s (so a timed block fills the
    GPU and each launch stays seconds long). Then per_stratum blocks of that
    size are drawn without replacement (all blocks when there are no more;
    the pilot blocks, whose draw decided the size, are not reused) and
    counted. Per stratum h with N_h blocks, n_h sampled, block sums y: total
    N_h * mean(y), variance N_h^2 (1 - n_h/N_h) var(y) / n_h (exact when
    n_h = N_h); device time likewise (an upper bound where blocks are a few
    units: a launch of few tasks runs its longest one alone). Returns
    {"raw": [...], "raw_se": [...], "seconds", "seconds_se", "unit",
    "strata": [...]} (raw sums are not divided by the divisor). With
    units=(lo, hi) only the units [lo, hi) are sampled and the sums are
    theirs alone."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n, unit, count, _ = unit_space(plan, config)
    u_lo, u_hi = units if units else (0, n)
    np_ = len(plan.patterns)
    raw, var = np.zeros(np_), np.zeros(np_)
    sec, sec_var = 0.0, 0.0
    strata = []

    def block(lo, hi, blk, b):
        r0 = lo + int(b) * blk
        r1 = min(hi, r0 + blk)
        got, st = count(r0, r1)
        if log and st["seconds"] > 1.0:
            log("  block [%d, %d) %s: %.2f s" % (r0, r1, unit, st["seconds"]))
        return got, st["seconds"]

    for lo, hi in strata_of(u_hi - u_lo):
        lo, hi = lo + u_lo, hi + u_lo
        size = hi - lo
        blk, pilot = 1, 0.0
        while True:  # pilot: the block size whose launch lasts min_block_s
            nb = -(-size // blk)
            b = int(rng.integers(nb))
            _, t = block(lo, hi, blk, b)
            pilot += t
            if t >= min_block_s or blk >= size:
                break
            blk = min(size, blk * (2 if t > min_block_s / 4 else 8))
        nb = -(-size // blk)
        pick = range(nb) if nb <= per_stratum else rng.choice(nb, per_stratum, replace=False)
        ys, ts = [], []
        for x in sorted(int(v) for v in pick):
            g2, t2 = block(lo, hi, blk, x)
            ys.append(g2)
            ts.append(t2)
        ys, ts = np.array(ys, dtype=np.float64), np.array(ts)
        k = len(ys)
        tot = nb * ys.mean(axis=0)
        fpc = 1.0 - k / nb
        v = (nb * nb * fpc * ys.var(axis=0, ddof=1) / k) if k > 1 and fpc > 0 else np.zeros(np_)
        tsum = nb * ts.mean()
        tv = nb * nb * fpc * ts.var(ddof=1) / k if k > 1 and fpc > 0 else 0.0
        raw += tot
        var += v
        sec += tsum
        sec_var += tv
        strata.append({"units": [lo, hi], "block": blk, "blocks": nb, "sampled": k,
                       "raw": tot.tolist(), "raw_se": np.sqrt(v).tolist(), "seconds": tsum,
                       "seconds_se": float(np.sqrt(tv)), "sampled_seconds": float(ts.sum()),
                       "pilot_seconds": pilot})
        if log:
            log("stratum [%d, %d) %s: %d of %d blocks of %d, %.3f s sampled, projected %.2f +- %.2f s"
                % (lo, hi, unit, k, nb, blk, ts.sum(), tsum, np.sqrt(tv)))
    return {"raw": raw.tolist(), "raw_se": np.sqrt(var).tolist(), "seconds": sec,
            "seconds_se": float(np.sqrt(sec_var)), "unit": unit, "strata": strata}


def combine(parts):
    """Counts from the --roots outputs of several calls (or GPUs) over one
    graph and pattern list: the root ranges must partition [0, vertices)
    (in any order), then the raw sums are added, divided by each pattern's
    divisor and multiplied by its multiplier, as gat_count reports them.
    Returns {"counts", "raw", "device_seconds", "roots"}."""
    if not parts:
        raise ValueError("no parts to combine")
    first = parts[0]
    for q in parts[1:]:
        for k in ("graph", "vertices", "patterns", "mode"):
            if q.get(k) != first.get(k):
                raise ValueError("parts differ in %s: %r vs %r" % (k, q.get(k), first.get(k)))
    ranges = sorted(tuple(q["roots"]) for q in parts)
    if ranges[0][0] != 0 or ranges[-1][1] != first["vertices"] or \
            any(a[1] != b[0] for a, b in zip(ranges, ranges[1:])):
        raise ValueError("root ranges %s do not partition [0, %d)" % (ranges, first["vertices"]))
    raw = [sum(int(q["raw"][i]) for q in parts) for i in range(len(first["raw"]))]
    counts = []
    for r, info in zip(raw, first["count_info"]):
        if r % info["divisor"]:
            raise ValueError("raw sum %d is not a multiple of the divisor %d" % (r, info["divisor"]))
        counts.append(r // info["divisor"] * info["multiplier"])
    return {"counts": counts, "raw": raw, "device_seconds": sum(q["device_seconds"] for q in parts),
            "roots": [list(r) for r in ranges]}


def make_graph(spec):
    """A workload graph by name (workloads.graph: erN, rmatS with seed S,
    clN) or a path to a .bin graph."""
    try:
        return workloads.graph(spec)
    except ValueError:
        return core.GlobalMemoryGraph.load(spec)


def main(argv=None):
    if argv is None:
        argv = sys.argv[1:]
    if argv and argv[0] == "--combine":
        # python -m graphautotune_amd.sharded --combine a.json b.json ...
        parts = []
        for path in argv[1:]:
            with open(path) as f:
                parts.append(json.loads(f.read().strip().splitlines()[-1]))
        print(json.dumps(combine(parts)), flush=True)
        return
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0],
                                 epilog="--combine A.json B.json ...: counts from the --roots outputs of a "
                                        "partition of the roots (see combine)")
    ap.add_argument("--graph", required=True, help="rmatS, clN or a .bin file")
    ap.add_argument("--pattern", action="append", default=[],
                    help="adjacency string or one of %s (repeatable)" % ", ".join(sorted(PATTERNS)))
    ap.add_argument("--batched-set", action="store_true",
                    help="add the batched configuration's 31 patterns (workloads.batched_patterns)")
    ap.add_argument("--mode", default="unique", choices=["unique", "labeled", "reference"])
    ap.add_argument("--config", default=None, help="JSON kernel config")
    ap.add_argument("--schedule-ranks", default=None, metavar="JSON",
                    help="a result of tuning.batched_study: run its per-pattern Schedule ranks (ranked "
                         "policy); its pattern list must equal the patterns given")
    ap.add_argument("--target", type=float, default=5.0, help="seconds per shard")
    ap.add_argument("--estimate", action="store_true",
                    help="project the whole-graph counts and device time from a stratified sample of root "
                         "blocks (see stratified_estimate) instead of counting every root")
    ap.add_argument("--per-stratum", type=int, default=8, help="with --estimate: blocks sampled per stratum")
    ap.add_argument("--min-block", type=float, default=1.0,
                    help="with --estimate: seconds a sampled block should take (see stratified_estimate)")
    ap.add_argument("--seed", type=int, default=0, help="with --estimate: sampling seed")
    ap.add_argument("--head-limit", type=float, default=None, metavar="S",
                    help="with --estimate: first time single units of roots 65536, 32768, ..., 0 (head_ladder), "
                         "stop at the first that takes S seconds or more, and sample only the units from the "
                         "last root under S on; the roots before it are reported as excluded")
    ap.add_argument("--roots", default=None, metavar="LO:HI",
                    help="count only roots [LO, HI) and report raw sums (add the raw sums of a partition, "
                         "then divide by the divisor and multiply by the multiplier); with --estimate, "
                         "sample only those roots' units")
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    g = make_graph(args.graph)
    log("graph %s: %d vertices, %d edges (%.1f s)" % (args.graph, g.v_cnt(), g.undirected_edges(),
                                                   time.perf_counter() - t0))
    pats = [PATTERNS.get(p, p) for p in args.pattern] + (workloads.batched_patterns() if args.batched_set else [])
    if not pats:
        ap.error("no patterns: give --pattern or --batched-set")
    sched = {}
    if args.schedule_ranks:
        with open(args.schedule_ranks) as f:
            tuned = json.load(f)
        if [core.Pattern(p).adj for p in tuned["patterns"]] != [core.Pattern(p).adj for p in pats]:
            ap.error("%s holds Schedule ranks for other patterns" % args.schedule_ranks)
        sched = {"schedule": "ranked", "schedule_rank": list(tuned["ranks"])}
    plan = core.Plan(g, pats, args.mode, **sched)
    cfg = json.loads(args.config) if args.config else {}
    if args.estimate:
        head, probes = 0, []
        n_units, _, _, first_unit = unit_space(plan, cfg)
        r_hi = plan.graph.v_cnt()
        if args.roots:  # only the units of the roots [LO, HI)
            head, r_hi = (int(x) for x in args.roots.split(":"))
            n_units = first_unit(r_hi) if r_hi < plan.graph.v_cnt() else n_units
        elif args.head_limit is not None:
            probes, head = head_ladder(plan, cfg, args.head_limit, log=log)
        est = stratified_estimate(plan, cfg, per_stratum=args.per_stratum, min_block_s=args.min_block,
                                  seed=args.seed, log=log, units=(first_unit(head), n_units))
        infos = [plan.count_info(i) for i in range(len(pats))]
        E = g.undirected_edges()
        print(json.dumps({
            "graph": args.graph, "vertices": g.v_cnt(), "undirected_edges": E, "patterns": pats, "mode": args.mode,
            "kernel_config": cfg, "schedule_ranks": sched.get("schedule_rank"), "per_stratum": args.per_stratum,
            "min_block_s": args.min_block, "seed": args.seed, "head_limit_s": args.head_limit,
            "head_probes": probes, "roots": [head, r_hi], "excluded_roots": [0, head] if not args.roots else None,
            "raw": est["raw"], "raw_se": est["raw_se"],
            "excluded_units": [0, first_unit(head)], "units": n_units, "unit": est["unit"],
            "estimated_counts": [r / i["divisor"] * i["multiplier"] for r, i in zip(est["raw"], infos)],
            "estimated_counts_se": [r / i["divisor"] * i["multiplier"] for r, i in zip(est["raw_se"], infos)],
            "projected_device_seconds": est["seconds"], "projected_device_seconds_se": est["seconds_se"],
            "projected_edges_per_s": E / est["seconds"] if est["seconds"] > 0 and head == 0 and not args.roots
            else None,
            "projected_units_per_s": (n_units - first_unit(head)) / est["seconds"] if est["seconds"] > 0 else None,
            "sampled_device_seconds": sum(st["sampled_seconds"] for st in est["strata"]),
            "strata": est["strata"]}), flush=True)
        return
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
