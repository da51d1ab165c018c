"""Manipulator: picks the next configs to measure by simulated annealing
over the cost model's predictions (tuning/tuning/role/manipulator.py:63-131:
40 iterations, T0 = 0.1, T *= 0.9, keep the batch_size best predictions)."""
import heapq
import math
import random

import numpy as np

from .driver import FLOAT_INF


class Manipulator:
    def __init__(self, model, driver, space, seed=0):
        self.model, self.driver, self.space = model, driver, space
        self.rng = random.Random(seed)
        self.np_rng = np.random.default_rng(seed)

    def random_batch(self, k):
        return self.space.random_configs(k, self.rng)

    def next_batch(self, batch_size, n_iter=40):
        temp = 0.1
        points = [self.space.random(self.rng) for _ in range(batch_size)]
        scores = np.asarray(self.model.predict_list(points), dtype=np.float64)
        # max-heap on predicted time of the current best batch (stored negated)
        heap = [(-s, i, p) for i, (s, p) in enumerate(zip(scores, points))]
        heapq.heapify(heap)
        in_heap = {p.fingerprint() for p in points}
        counter = len(points)
        for _ in range(n_iter):
            new = [self.space.random_walk(p, self.rng) for p in points]
            new_scores = np.asarray(self.model.predict_list(new), dtype=np.float64)
            with np.errstate(over="ignore"):
                ac_prob = np.exp(np.clip((scores - new_scores) / temp, -700, 700))
            accept = self.np_rng.random(len(ac_prob)) < ac_prob
            for i in range(len(points)):
                if accept[i]:
                    points[i], scores[i] = new[i], new_scores[i]
            for s, p in zip(new_scores, new):
                fp = p.fingerprint()
                if s < -heap[0][0] and fp not in in_heap and not self.model.is_config_exist(p):
                    _, _, old = heapq.heapreplace(heap, (-s, counter, p))
                    counter += 1
                    in_heap.discard(old.fingerprint())
                    in_heap.add(fp)
            temp *= 0.9
        return [p for _, _, p in sorted(heap, key=lambda t: -t[0])]

    def run_batch(self, trials):
        times = self.driver.batch_run(trials)
        ok = [(t, y) for t, y in zip(trials, times) if y != FLOAT_INF]
        return [t for t, _ in ok], [y for _, y in ok]

    def update_model(self, xs, ys):
        if xs:
            self.model.update_list(xs, ys)
