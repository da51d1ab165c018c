"""Learned config ranking at the sizes of the learned-tuning workload: the
MLP cost model is trained on runtimes the engine measures for random kernel
configs on R-MAT scales 16-20, then ranks a pool of configs for the batched
set's 6-vertex patterns on R-MAT scale 21. Only the model's top `budget`
picks count as its answer; the whole pool is measured too, as the grid
judge, and the model is compared with picking `budget` configs at random
(the method without it) and with the default config.

Each (graph, pattern) workload is a fixed root shard (gat_count_shard) of
mid-degree roots, grown from one root until the default config takes
`target_ms` (a workload whose one root takes over 4 target_ms is left out),
so that a 6-vertex pattern on a scale-21 graph is measured in milliseconds
rather than hours; every config of a workload must give the same raw sums. A
config whose count of the shard's first eighth already takes 2.5 times the
default's whole shard is scored by that time scaled to the shard (Driver
probe): it is far from the best either way, and its full run could take
minutes. The model regresses log(runtime / default runtime) of its
workload, which is what ranking within a workload needs.

    python -m graphautotune_amd.tuning.config_study --out results/config_study_rmat16-20_to_rmat21.json
    python -m graphautotune_amd.tuning.config_study --judge-clique --out results/clique_grid_rmat20.json

Reference: tuning/tuning/role/modeling.py:81-102,152-162 (fit on measured
runtimes, predict to rank) and tuning/tuning/tuner.py:28-53 (measure the
picks); XGBoost is replaced by the HIP MLP.
"""
import argparse
import json
import random
import sys
import time

import numpy as np

from .. import core, workloads
from . import features
from .driver import FLOAT_INF, Driver
from .mlp import MLPCostModel
from .space import Config, ConfigSpace

KNOBS = ["block_waves", "waves_per_cu", "queue_chunk", "slot_capacity", "intersect_strategy", "gallop_ratio",
         "search_linear_threshold"]
# always-linear scans of hub lists (set_search_type 64 and INT_MAX) are
# pathological; the array store is the engine these patterns run on
FIXED = {"vertex_store": 1}
DROP = {"search_linear_threshold": [64, 2 ** 31 - 1]}


def space():
    sp = ConfigSpace(KNOBS, fixed=FIXED)
    for i, p in enumerate(sp.params):
        if p.name in DROP:
            import copy
            q = copy.copy(p)
            q.values = [v for v in p.values if v not in DROP[p.name]]
            sp.params[i] = q
    return sp


def rmat(scale):
    return core.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, scale)


def grow_shard(plan, lo, n, target_ms, first=1):
    """Roots [lo, lo + size): size doubled (or multiplied by 8 while far
    below target) from `first` until the default config takes target_ms or
    the shard reaches the last root; returns (hi, default seconds)."""
    base = dict(FIXED)
    size = first
    while True:
        hi = min(n, lo + size)
        plan.count_shard(lo, hi, base)
        _, st = plan.count_shard(lo, hi, base)
        ms = st["seconds"] * 1e3
        if ms >= target_ms or hi == n:
            return hi, st["seconds"]
        size *= 2 if ms > target_ms / 8 else 8


def pick_shard(graph, pattern, target_ms, first=1, log=None):
    """A root shard [lo, hi) of the degree-ordered graph whose default-config
    count takes about target_ms (grow_shard), starting at lo = n / 16
    (mid-degree roots). When every root from lo on together takes under a
    quarter of the target, lo moves to heavier roots (lo / 4); when one root
    alone takes over four times the target, to lighter ones (4 lo).
    Returns (lo, hi, default seconds)."""
    n = graph.v_cnt()
    plan = core.Plan(graph, [pattern], "unique")
    lo, tried = n // 16, set()
    while True:
        tried.add(lo)
        hi, t = grow_shard(plan, lo, n, target_ms, first)
        ms = t * 1e3
        nxt = lo
        if ms < target_ms / 4 and hi == n and lo > 0:
            nxt = lo // 4
        elif ms > 4 * target_ms and hi - lo == first and lo < n - 1:
            nxt = min(n - 1, 4 * lo)
        if nxt == lo or nxt in tried:
            if log:
                log("  shard [%d, %d): default %.2f ms" % (lo, hi, ms))
            return lo, hi, t
        lo = nxt


def featurize(graph, pattern):
    plan = core.Plan(graph, [pattern], "unique")
    pf = features.pattern_features(plan.schedule(0), 1)
    gf = features.graph_features(graph)
    return lambda cfg: features.vector(pf, gf, cfg)


def default_config():
    return Config(dict(FIXED))


def collect(scales, patterns_per_scale, train_pool, per_workload, target_ms, seed=0, log=print):
    """Random configs measured on root shards of R-MAT graphs; returns the
    feature rows, runtime / default runtime, and per-run records."""
    rng = random.Random(seed)
    sp = space()
    X, y, meta = [], [], []
    for sc in scales:
        g = rmat(sc)
        pats = rng.sample(train_pool, patterns_per_scale)
        for p in pats:
            t0 = time.time()
            lo, hi, tp = pick_shard(g, p, target_ms, log=log)
            if tp * 1e3 < target_ms / 4:
                log("collect s%d %s: %.2f ms on all roots from %d, too small to time; left out" % (sc, p, tp * 1e3, lo))
                continue
            if tp * 1e3 > 4 * target_ms:
                log("collect s%d %s: %.2f ms on one root, too heavy to time many configs; left out" % (sc, p, tp * 1e3))
                continue
            drv = Driver(g, [p], "unique", reps=1, warmup=1, roots=(lo, hi))
            td = drv.run(default_config())
            drv.probe = (max(1, (hi - lo) // 8), 2.5 * td)
            drv.report = log
            feat = featurize(g, p)
            for c in [default_config()] + sp.random_configs(per_workload, rng):
                t = td if c == default_config() else drv.run(c)
                if t >= FLOAT_INF:
                    continue
                X.append(feat(c))
                y.append(t / td)
                meta.append({"scale": sc, "pattern": p, "roots": [lo, hi], "config": c.as_dict(), "seconds": t})
            log("collect s%d %s roots [%d, %d): default %.2f ms, %d configs in %.1f s"
                % (sc, p, lo, hi, td * 1e3, per_workload, time.time() - t0))
    return np.stack(X), np.asarray(y), meta


def expected_min_of_random(values, k):
    """E[min of k distinct picks drawn uniformly from values] (exact)."""
    v = np.sort(np.asarray(values, dtype=np.float64))
    n = len(v)
    k = min(k, n)
    # the i-th smallest (0-based) is the minimum of a k-subset with
    # probability C(n-1-i, k-1) / C(n, k)
    from math import comb
    tot = comb(n, k)
    return float(sum(v[i] * comb(n - 1 - i, k - 1) for i in range(n - k + 1)) / tot)


def rank(model, scale, patterns, pool_size, budget, target_ms, seed=0, log=print, reps=2):
    """Per pattern: measure a seeded pool of configs plus the default (the
    grid; median of `reps` runs each), rank the pool with the model, take its
    top `budget`; compare with `budget` random picks (exact expectation) and
    the default config. A pattern whose whole graph takes less than
    target_ms / 4 with the default config is too small to time and is left
    out (listed in the log)."""
    rng = random.Random(seed + 1)
    sp = space()
    g = rmat(scale)
    out = []
    for p in patterns:
        lo, hi, t0 = pick_shard(g, p, target_ms, log=log)
        pool = [default_config()] + sp.random_configs(pool_size, rng)
        if t0 * 1e3 < target_ms / 4:
            log("rank s%d %s: %.2f ms on all roots from %d, too small to time; left out" % (scale, p, t0 * 1e3, lo))
            continue
        if t0 * 1e3 > 4 * target_ms:
            log("rank s%d %s: %.2f ms on one root, too heavy to time many configs; left out" % (scale, p, t0 * 1e3))
            continue
        drv = Driver(g, [p], "unique", reps=reps, warmup=1, roots=(lo, hi))
        td = drv.run(default_config())
        drv.probe = (max(1, (hi - lo) // 8), 2.5 * td)
        drv.report = log
        truth = np.array([td] + [drv.run(c) for c in pool[1:]])
        ok = truth < FLOAT_INF
        feat = featurize(g, p)
        pred = model.predict(np.stack([feat(c) for c in pool]))
        order = [i for i in np.argsort(pred) if ok[i]]
        picks = order[:budget]
        best = truth[ok].min()
        res = {
            "pattern": p, "scale": scale, "roots": [lo, hi], "pool": int(ok.sum()), "budget": budget,
            "grid_best_ms": best * 1e3, "default_ms": td * 1e3,
            "model_pick_ms": truth[picks].min() * 1e3, "model_picks": [int(i) for i in picks],
            "model_regret": truth[picks].min() / best - 1.0,
            "random_regret": expected_min_of_random(truth[ok], budget) / best - 1.0,
            "default_regret": td / best - 1.0,
            "spearman": spearman(pred[ok], truth[ok]),
            "counts": drv.counts,
        }
        log(json.dumps(res))
        out.append(res)
    return out


def spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


def run_study(train_scales, eval_scale, patterns_per_scale, per_workload, pool, budget, target_ms, steps, seed,
              log=print):
    held = workloads.batched_patterns()[21:]  # the batched set's 6-vertex patterns
    six = [p for p in workloads.connected_patterns(6) if p not in held]
    train_pool = list(workloads.connected_patterns(5)) + six
    t0 = time.time()
    X, y, meta = collect(train_scales, patterns_per_scale, train_pool, per_workload, target_ms, seed, log)
    t_collect = time.time() - t0
    model = MLPCostModel(features.FEATURES, 256, seed=seed)
    loss = model.fit(X, y, steps=steps, lr=1e-3)
    log("trained on %d runs: loss %.4f" % (len(y), loss))
    t0 = time.time()
    res = rank(model, eval_scale, held, pool, budget, target_ms, seed, log)
    return {
        "train": {"scales": list(train_scales), "runs": len(y), "workloads": len(set((m["scale"], m["pattern"])
                                                                                  for m in meta)),
                  "loss": loss, "collect_seconds": t_collect},
        "eval": {"scale": eval_scale, "patterns": len(res), "held_out": len(held), "pool": pool + 1, "budget": budget,
                 "seconds": time.time() - t0},
        "mean_model_regret": float(np.mean([r["model_regret"] for r in res])),
        "mean_random_regret": float(np.mean([r["random_regret"] for r in res])),
        "mean_default_regret": float(np.mean([r["default_regret"] for r in res])),
        "mean_spearman": float(np.mean([r["spearman"] for r in res])),
        "sum_grid_best_ms": float(sum(r["grid_best_ms"] for r in res)),
        "sum_model_pick_ms": float(sum(r["model_pick_ms"] for r in res)),
        "sum_default_ms": float(sum(r["default_ms"] for r in res)),
        "results": res,
        "train_runs": meta,
    }


def judge_clique_config(scale=20, knobs=("queue_chunk", "slot_capacity", "block_waves", "waves_per_cu"), log=print):
    """The benchmark's tuned K4 + K5 config (tuned_configs.json) against a
    grid over `knobs` with the other knobs at their tuned values, on the
    benchmark graph: the grid is the judge of the tuner's pick."""
    import os
    from .space import BY_NAME
    from .tuner import grid_search
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_configs.json")
    with open(path) as f:
        tuned = json.load(f)["rmat%d-k4k5" % scale]["config"]
    g = core.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, scale)
    pats = [workloads.K4, workloads.K5]
    drv = Driver(g, pats, "unique", reps=3, warmup=1)
    t_tuned = drv.run(Config(tuned))
    fixed = {k: v for k, v in tuned.items() if k not in knobs}
    best, t_best, results = grid_search(drv, ConfigSpace(list(knobs), fixed=fixed), log=log)
    ranks = sorted(t for _, t in results)
    return {"graph": "rmat%d" % scale, "patterns": pats, "knobs": list(knobs),
            "grid_points": len(results), "grid_values": {k: BY_NAME[k].values for k in knobs},
            "tuned_config": tuned, "tuned_ms": t_tuned * 1e3, "grid_best_config": best.as_dict(),
            "grid_best_ms": t_best * 1e3, "tuned_regret": t_tuned / t_best - 1.0,
            "tuned_rank": sum(1 for t in ranks if t < t_tuned) + 1, "counts": drv.counts,
            "grid": [[c.as_dict(), t * 1e3] for c, t in results]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--train-scales", default="16,17,18,19,20")
    ap.add_argument("--eval-scale", type=int, default=21)
    ap.add_argument("--patterns-per-scale", type=int, default=6)
    ap.add_argument("--per-workload", type=int, default=12, help="random configs measured per training workload")
    ap.add_argument("--pool", type=int, default=24, help="random configs per evaluation pattern (plus the default)")
    ap.add_argument("--budget", type=int, default=3, help="picks the model (or random) may measure")
    ap.add_argument("--target-ms", type=float, default=40.0, help="default-config time of a workload's shard")
    ap.add_argument("--steps", type=int, default=800)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--judge-clique", action="store_true",
                    help="instead: time the benchmark's tuned K4 + K5 config against a grid of queue chunk x "
                         "slot capacity x block waves x waves per CU on R-MAT scale 20")
    args = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if args.judge_clique:
        res = judge_clique_config(log=log)
        res["command"] = "python -m graphautotune_amd.tuning.config_study " + " ".join(argv or sys.argv[1:])
        print(json.dumps({k: v for k, v in res.items() if k != "grid"}), flush=True)
        if args.out:
            with open(args.out, "w") as f:
                json.dump(res, f, indent=1)
        return
    summary = run_study([int(s) for s in args.train_scales.split(",")], args.eval_scale, args.patterns_per_scale,
                        args.per_workload, args.pool, args.budget, args.target_ms, args.steps, args.seed, log)
    summary["command"] = "python -m graphautotune_amd.tuning.config_study " + " ".join(argv or sys.argv[1:])
    print(json.dumps({k: v for k, v in summary.items() if k not in ("results", "train_runs")}), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(summary, f, indent=1)


if __name__ == "__main__":
    main()
