"""Tuner: warm up on random configs, then rounds of model-guided batches
(tuning/tuning/role/tuner.py:23-53). Grid search is the fallback and the
judge of the model's picks."""
from .driver import FLOAT_INF


class Tuner:
    def __init__(self, manipulator, max_round=20, batch_size=5, warmup_examples=20):
        self.m = manipulator
        self.max_round, self.batch_size, self.warmup_examples = max_round, batch_size, warmup_examples

    def tune(self, log=None):
        xs, ys = self.m.run_batch(self.m.random_batch(self.warmup_examples))
        self.m.update_model(xs, ys)
        for r in range(self.max_round):
            xs, ys = self.m.run_batch(self.m.next_batch(self.batch_size))
            self.m.update_model(xs, ys)
            best = self.m.model.get_best_config()
            if log:
                log("round %d/%d best %.4f ms %s" % (r + 1, self.max_round, best[1] * 1e3, best[0].as_dict()))
        return self.m.model.get_best_config()


def grid_search(driver, space, log=None):
    best, best_t, results = None, FLOAT_INF, []
    for c in space.all():
        t = driver.run(c)
        results.append((c, t))
        if t < best_t:
            best, best_t = c, t
            if log:
                log("grid best %.4f ms %s" % (t * 1e3, c.as_dict()))
    return best, best_t, results
