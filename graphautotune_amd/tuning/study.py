"""The learned-tuning study: collect engine-measured (config, runtime) pairs
on small R-MAT graphs, train the MLP cost model on the GPU, then let it rank
unseen configs for 6-vertex patterns on a larger R-MAT graph. The ranking is
judged against measured runtimes of the whole candidate pool (the grid),
and compared with picking at random under the same measurement budget."""
import json
import random
import time

import numpy as np

from .. import core
from . import features
from .driver import FLOAT_INF, Driver
from .mlp import MLPCostModel
from .space import Config, ConfigSpace

TUNE_KNOBS = ["block_waves", "waves_per_cu", "queue_chunk", "slot_capacity", "intersect_strategy",
              "gallop_ratio", "search_linear_threshold"]
# Linear scans of 64+ entries (the reference's set_search_type 64 and
# INT_MAX, i.e. always linear) are pathological on hub neighbour lists;
# studies and tuning sessions search the rest of the space.
STUDY_FIXED_OUT = {"search_linear_threshold": [64, 2 ** 31 - 1]}


def study_space():
    sp = ConfigSpace(TUNE_KNOBS)
    for p in sp.params:
        if p.name in STUDY_FIXED_OUT:
            import copy
            q = copy.copy(p)
            q.values = [v for v in p.values if v not in STUDY_FIXED_OUT[p.name]]
            sp.params[sp.params.index(p)] = q
    return sp


def featurizer(graph, patterns):
    plan = core.Plan(graph, patterns, "unique")
    sched = max((plan.schedule(i) for i in range(len(patterns))), key=lambda s: s.data.v_cnt)
    pf = features.pattern_features(sched, len(patterns))
    gf = features.graph_features(graph)
    return lambda cfg: features.vector(pf, gf, cfg)


def collect(scales, patterns, per_workload, seed=0, log=print, reps=1, max_default_s=0.25):
    """Measure random configs on each (scale, pattern); returns X, y, meta.
    A workload whose default config takes longer than max_default_s is
    skipped so the study stays within a bounded GPU time."""
    rng = random.Random(seed)
    space = study_space()
    X, y, meta = [], [], []
    for sc in scales:
        g = core.Graph.rmat(sc, 16, 0.57, 0.19, 0.19, sc)
        for p in patterns:
            drv = Driver(g, [p], "unique", reps=reps, warmup=0)
            feat = featurizer(g, [p])
            t0 = time.time()
            td = drv.run(Config())
            log("collect s%d %s: default %.2f ms (counts %s)" % (sc, p, td * 1e3, drv.counts))
            if td > max_default_s:
                log("  skipped: default config above %.2f s" % max_default_s)
                continue
            for c in space.random_configs(per_workload, rng):
                t = drv.run(c)
                log("  %.3f ms %s" % (t * 1e3, c.as_dict()))
                if t < FLOAT_INF:
                    X.append(feat(c))
                    y.append(t)
                    meta.append({"scale": sc, "pattern": p, "config": c.as_dict(), "seconds": t})
            log("  %d configs in %.1fs" % (per_workload, time.time() - t0))
    return np.stack(X), np.asarray(y), meta


def spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


def rank_study(model, scale, patterns, pool_size, budget, seed=0, log=print, max_default_s=0.5):
    """For each pattern: measure every config of a random pool (the grid
    judge), rank the pool with the model, re-measure the model's top-`budget`
    picks, and compare with `budget` random picks."""
    rng = random.Random(seed + 1)
    space = study_space()
    g = core.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, scale)
    out = []
    for p in patterns:
        drv = Driver(g, [p], "unique", reps=1, warmup=0)
        feat = featurizer(g, [p])
        td = drv.run(Config())
        log("rank s%d %s: default %.2f ms" % (scale, p, td * 1e3))
        if td > max_default_s:
            log("  skipped: default config above %.2f s" % max_default_s)
            continue
        pool = space.random_configs(pool_size, rng)
        truth = []
        for c in pool:
            truth.append(drv.run(c))
            log("  pool %.3f ms" % (truth[-1] * 1e3))
        truth = np.array(truth)
        ok = truth < FLOAT_INF
        pred = model.predict(np.stack([feat(c) for c in pool]))
        best = truth[ok].min()
        order = np.argsort(pred)
        picks = [i for i in order if ok[i]][:budget]
        model_best = truth[picks].min()
        rnd = []
        for trial in range(200):
            sample = rng.sample([i for i in range(pool_size) if ok[i]], budget)
            rnd.append(truth[sample].min())
        res = {
            "pattern": p, "scale": scale, "pool": int(ok.sum()), "budget": budget,
            "grid_best_ms": best * 1e3,
            "model_pick_ms": model_best * 1e3,
            "model_regret": model_best / best - 1.0,
            "random_pick_ms_mean": float(np.mean(rnd)) * 1e3,
            "random_regret_mean": float(np.mean(rnd)) / best - 1.0,
            "spearman": spearman(pred[ok], truth[ok]),
            "default_ms": td * 1e3,
            "counts": drv.counts,
        }
        log(json.dumps(res))
        out.append(res)
    return out
