"""Per-pattern Schedule choice by measurement.

The reference builds one Schedule per pattern from a cost estimate fed with
fixed graph statistics (core/schedule.hpp:909-910) and never revisits it.
Every candidate Schedule counts the same instances in unique and labeled
mode, but their running times differ by large factors on skewed graphs. This
module times the candidates of the "ranked" policy (ranked under the
reference's fixed statistics, so a rank names the same Schedule on every
graph) on a proxy graph, single pattern at a time, and keeps the fastest per
pattern; the chosen ranks then go to a batched plan on the target graph.
Rank 0 (the reference's pick) is always a candidate, so the choice is never
worse on the proxy.
"""
import time

from ..core import Plan, schedule_candidates


def tune_schedule_ranks(proxy, patterns, mode="unique", max_rank=12, config=None, reps=1, log=None):
    """Returns (ranks, per-pattern [(rank, seconds)] measurements)."""
    ranks, history = [], []
    for i, p in enumerate(patterns):
        _, total = schedule_candidates(p, limit=1)
        results = []
        base = None
        for r in range(min(max_rank, total)):
            plan = Plan(proxy, [p], mode, schedule="ranked", schedule_rank=r)
            plan.count(config)  # warm-up (device buffers)
            best = None
            for _ in range(reps):
                counts, st = plan.count(config)
                best = st["seconds"] if best is None else min(best, st["seconds"])
            if base is None:
                base = counts
            elif counts != base:
                raise AssertionError("schedule rank %d of pattern %d changed the count: %s vs %s"
                                     % (r, i, counts, base))
            results.append((r, best))
        r_best, t_best = min(results, key=lambda x: (x[1], x[0]))
        ranks.append(r_best)
        history.append(results)
        if log:
            log("pattern %d: %d candidates, rank 0 %.2f ms, best rank %d %.2f ms"
                % (i, total, results[0][1] * 1e3, r_best, t_best * 1e3))
    return ranks, history
