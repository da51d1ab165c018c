"""Cost-model inputs: pattern adjacency (upper triangle of 7x7, schedule
order), schedule loop order, a 32-bin log-degree histogram of the graph and
one-hot kernel-config knobs, padded to FEATURES floats."""
import math

import numpy as np

from .space import SPACE

FEATURES = 128
MAXN = 7


def pattern_features(schedule, n_patterns=1):
    d = schedule.data
    n = d.v_cnt
    f = np.zeros(31, dtype=np.float32)
    k = 0
    for i in range(MAXN):
        for j in range(i + 1, MAXN):
            if i < n and j < n and (d.adj_rows[i] >> j) & 1:
                f[k] = 1.0
            k += 1
    for i in range(min(n, MAXN)):
        f[21 + i] = (schedule.perm[i] + 1) / MAXN
    f[28] = schedule.basic_vertexes / MAXN
    f[29] = schedule.iep_suffix_vertexes / MAXN
    f[30] = min(n_patterns, 32) / 32.0
    return f


def graph_features(graph):
    h = graph.degree_histogram().astype(np.float64)
    v = max(1, graph.v_cnt())
    f = np.zeros(35, dtype=np.float32)
    f[:32] = h / v
    f[32] = math.log2(v) / 32
    f[33] = math.log2(max(1, graph.undirected_edges())) / 40
    f[34] = math.log2(max(1, graph.max_degree())) / 32
    return f


def config_features(config):
    parts = []
    for p in SPACE:
        oh = np.zeros(len(p.values), dtype=np.float32)
        oh[p.values.index(config.values[p.name])] = 1.0
        parts.append(oh)
    return np.concatenate(parts)


def vector(pattern_f, graph_f, config):
    x = np.zeros(FEATURES, dtype=np.float32)
    c = config_features(config)
    body = np.concatenate([pattern_f, graph_f, c])
    assert len(body) <= FEATURES, len(body)
    x[:len(body)] = body
    return x
