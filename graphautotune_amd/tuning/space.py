"""Discrete kernel-config space (tuning/tuning/config/details.py:11-85 and
base.py:93-342 of the reference, for the runtime knobs of gat_kernel_config).

The reference's Config rebuilds the C++ code per point; here a point is a
dict handed to gat_count. The index vector sorted by key (base.py:218-222)
and its MD5 fingerprint (base.py:259-262) are kept.
"""
import hashlib
import itertools
import random as _random

INT_MAX = 2 ** 31 - 1


class Param:
    def __init__(self, name, values, default, ref):
        self.name, self.values, self.default, self.ref = name, list(values), default, ref
        assert default in self.values, (name, default)


SPACE = [
    Param("block_waves", [1, 2, 4], 1, "ThreadsPerBlock (details.py:50-52) / 64"),
    Param("waves_per_cu", [4, 8, 12, 16], 8, "NumBlocks (details.py:45-47), per CU"),
    Param("queue_chunk", [1, 4, 16, 64], 16, "edges per work-queue grab (prefix chunk)"),
    Param("slot_capacity", [32, 64, 128, 256, 512, 1024], 1024,
          "LDS set-buffer entries (array store); per-wave row limit d <= min(256, value) (bitmap store)"),
    Param("intersect_strategy", [0, 1, 2, 3, 4], 0, "auto / stream-and-search / stream-shorter / merge path / galloping"),
    Param("gallop_ratio", [4, 32, 256], 32, "auto: skew ratio above which the slow side is not streamed"),
    Param("search_linear_threshold", [0, 4, 16, 64, INT_MAX], 0, "SetSearchType (details.py:11-13)"),
    Param("vertex_store", [0, 1], 0, "VertexStoreType (details.py:21-23): auto(bitmap for cliques) / array"),
]
BY_NAME = {p.name: p for p in SPACE}


class Config:
    """One point of the space (a dict of knob -> value)."""

    def __init__(self, values=None):
        self.values = {p.name: p.default for p in SPACE}
        for k, v in (values or {}).items():
            if k not in BY_NAME:
                raise KeyError("unknown knob %r" % k)
            if v not in BY_NAME[k].values:
                raise ValueError("invalid value %r for %s; valid: %s" % (v, k, BY_NAME[k].values))
            self.values[k] = v

    def copy(self):
        return Config(dict(self.values))

    def get_list(self):
        return [BY_NAME[k].values.index(self.values[k]) for k in sorted(self.values)]

    def fingerprint(self):
        return hashlib.md5(",".join(str(x) for x in self.get_list()).encode()).hexdigest()

    def as_dict(self):
        return dict(self.values)

    def __eq__(self, other):
        return isinstance(other, Config) and self.values == other.values

    def __hash__(self):
        return hash(self.fingerprint())

    def __repr__(self):
        return "Config(%s)" % self.values


class ConfigSpace:
    def __init__(self, params=None, fixed=None):
        """params: knob names that vary (default: all); fixed: pinned values."""
        self.fixed = dict(fixed or {})
        self.params = [BY_NAME[n] for n in (params or [p.name for p in SPACE]) if n not in self.fixed]
        self.params = [p for p in self.params if len(p.values) > 1]

    def size(self):
        n = 1
        for p in self.params:
            n *= len(p.values)
        return n

    def all(self):
        for combo in itertools.product(*[p.values for p in self.params]):
            d = dict(self.fixed)
            d.update({p.name: v for p, v in zip(self.params, combo)})
            yield Config(d)

    def random(self, rng=_random):
        d = dict(self.fixed)
        d.update({p.name: rng.choice(p.values) for p in self.params})
        return Config(d)

    def random_configs(self, k, rng=_random):
        if k > self.size():
            raise ValueError("cannot sample %d configs from a space of size %d" % (k, self.size()))
        out, seen = [], set()
        while len(out) < k:
            c = self.random(rng)
            if c.fingerprint() not in seen:
                seen.add(c.fingerprint())
                out.append(c)
        return out

    def random_walk(self, config, rng=_random):
        """Change one knob to a different value (base.py:308-342)."""
        p = rng.choice(self.params)
        new = config.copy()
        cur = new.values[p.name]
        new.values[p.name] = rng.choice([v for v in p.values if v != cur])
        return new
