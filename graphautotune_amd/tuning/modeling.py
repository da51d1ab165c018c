"""Modeling: history of (config, runtime) plus the learned cost model
(tuning/tuning/role/modeling.py:16-162; XGBoost replaced by the HIP MLP).
The history is a JSON file rewritten after every update (modeling.py:70-79)
and reloaded at start (:48-68)."""
import json
import os

import numpy as np

from . import features
from .mlp import MLPCostModel
from .space import Config

FLOAT_INF = 1e38


class Modeling:
    def __init__(self, featurize, record_path=None, seed=0, steps=400, lr=2e-3):
        """featurize(Config) -> feature row for this workload."""
        self.featurize = featurize
        self.record_path = record_path
        self.xs, self.ys = [], []
        self.model = MLPCostModel(features.FEATURES, 256, seed)
        self.steps, self.lr = steps, lr
        self._load()
        self._fit()

    def _load(self):
        if not self.record_path or not os.path.exists(self.record_path) or os.path.getsize(self.record_path) == 0:
            return
        with open(self.record_path) as f:
            for cfg, y in json.load(f):
                self.xs.append(Config(cfg))
                self.ys.append(float(y))

    def _dump(self):
        if self.record_path:
            with open(self.record_path, "w") as f:
                json.dump([[x.as_dict(), y] for x, y in zip(self.xs, self.ys)], f)

    def _fit(self):
        pts = [(x, y) for x, y in zip(self.xs, self.ys) if y < FLOAT_INF]
        if not pts:
            return
        X = np.stack([self.featurize(x) for x, _ in pts])
        self.model.fit(X, [y for _, y in pts], steps=self.steps, lr=self.lr)

    def is_config_exist(self, c):
        return c.fingerprint() in {x.fingerprint() for x in self.xs}

    def predict_list(self, configs):
        if not self.model.trained:
            return np.zeros(len(configs))
        return self.model.predict(np.stack([self.featurize(c) for c in configs]))

    def update_list(self, xs, ys):
        self.xs.extend(xs)
        self.ys.extend(ys)
        self._dump()
        self._fit()

    def get_best_config(self):
        if not self.xs:
            return Config(), FLOAT_INF
        i = int(np.argmin(self.ys))
        return self.xs[i], self.ys[i]
