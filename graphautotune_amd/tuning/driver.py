"""Driver: runs one kernel config of a workload and returns its time
(tuning/tuning/role/driver.py:13-132 compiles a binary per trial and reads
result files; here the engine runs in-process over the C ABI).

Every run's counts must equal the first run's: a config that changes a
count is a bug, not a fast config. An infeasible config (e.g. too much LDS)
returns FLOAT_INF, as the reference maps a failing run (driver.py:118-122).
An engine inconsistency (GAT_EINTERNAL) is raised as EngineBug, never
scored: it says nothing about the config's speed.
"""
import statistics

from .. import core
from .._lib import GAT_EINFEASIBLE, GAT_EINTERNAL, GatError

FLOAT_INF = 1e38  # tuning/tuning/common/const.py:3


class EngineBug(RuntimeError):
    """The engine reported an internal inconsistency for a config."""


class Driver:
    """roots=(lo, hi) times a root shard (gat_count_shard, raw sums) instead
    of the whole graph: on large graphs a fixed shard is the workload.
    plan_options are passed to core.Plan (e.g. schedule="ranked" with
    per-pattern schedule_rank). probe=(k, limit_s), with roots: each config
    first counts the shard's first k roots once, and when that takes limit_s
    or more the config is scored by that time scaled to the shard's size
    instead of being run whole (a config tens of times slower than the
    default tells a ranker all it needs, and its full run could take
    minutes); such scores are marked in the report."""

    def __init__(self, graph, patterns, mode="unique", reps=3, warmup=1, roots=None, plan_options=None,
                 report=None):
        self.plan = core.Plan(graph, patterns, mode, **(plan_options or {}))
        self.report = report  # called with a line per measured config
        self.graph = graph
        self.patterns = list(patterns)
        self.reps, self.warmup = reps, warmup
        self.roots = roots
        self.probe = None
        self.counts = None
        self.probe_counts = None
        self.log = []

    def _count(self, cfg):
        if self.roots:
            return self.plan.count_shard(self.roots[0], self.roots[1], cfg)
        return self.plan.count(cfg)

    def run(self, config):
        cfg = config.as_dict() if hasattr(config, "as_dict") else dict(config)
        try:
            if self.probe and self.roots and self.roots[1] - self.roots[0] > self.probe[0]:
                k, limit = self.probe
                lo, hi = self.roots
                c, st = self.plan.count_shard(lo, lo + k, cfg)
                if self.probe_counts is None:
                    self.probe_counts = c
                elif c != self.probe_counts:
                    raise AssertionError("config %s changed the counts: %s vs %s" % (cfg, c, self.probe_counts))
                if st["seconds"] >= limit:
                    t = st["seconds"] * (hi - lo) / k
                    self.log.append((cfg, t))
                    if self.report:
                        self.report("scaled from %d of %d roots: %.3f ms %s" % (k, hi - lo, t * 1e3, cfg))
                    return t
            for _ in range(self.warmup):
                self._count(cfg)
            secs = []
            for _ in range(self.reps):
                c, st = self._count(cfg)
                if self.counts is None:
                    self.counts = c
                elif c != self.counts:
                    raise AssertionError("config %s changed the counts: %s vs %s" % (cfg, c, self.counts))
                secs.append(st["seconds"])
        except GatError as e:
            if e.code == GAT_EINFEASIBLE:
                self.log.append((cfg, FLOAT_INF))
                return FLOAT_INF
            if e.code == GAT_EINTERNAL:
                raise EngineBug("engine bug under config %s: %s" % (cfg, e)) from e
            raise
        t = statistics.median(secs)
        self.log.append((cfg, t))
        if self.report:
            self.report("measured %.3f ms %s" % (t * 1e3, cfg))
        return t

    def batch_run(self, configs):
        return [self.run(c) for c in configs]
