"""Host-side mirror of the reference's operator interface for the `pm` path.

Names follow the reference so that code written against it reads the same:

    schedule = Schedule(Pattern(adj))                 # core/schedule.hpp:864
    graph = GlobalMemoryGraph.load(path, binary=True) # infra/graph_backend.cuh:155
    context = DeviceContext(schedule, graph)          # engine/context.cuh:18
    engine = Executor(graph.max_degree(), GPU_DEVICE) # engine/engine.cuh:157
    answer = engine.perform_search(context)           # engine/engine.cuh:232

`perform_search` returns the reference engine's value for the schedule
(mode "reference"); `count()` exposes all modes and many patterns at once.
Everything below the Python layer is libgat.so (host C++ + HIP for gfx950).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MODE_LABELED, MODE_REFERENCE, MODE_UNIQUE, SCHEDULE_GRAPH, SCHEDULE_RANKED, SCHEDULE_REFERENCE,
                   CountStats, KernelConfig, PlanOptions, ScheduleData, check, lib)

CPU_DEVICE, GPU_DEVICE, UNKNOWN_DEVICE = 0, 1, 2  # consts/general_consts.hpp:6-10

MODES = {"unique": MODE_UNIQUE, "reference": MODE_REFERENCE, "labeled": MODE_LABELED}


class Pattern:
    """n x n row-major '0'/'1' adjacency string (core/pattern.hpp:22-39)."""

    def __init__(self, adj):
        if isinstance(adj, Pattern):
            adj = adj.adj
        self.adj = str(adj)
        n = int(round(len(self.adj) ** 0.5))
        if n * n != len(self.adj) or set(self.adj) - {"0", "1"}:
            raise ValueError("pattern must be an n*n string of '0'/'1', got %r" % adj)
        self._n = n

    def v_cnt(self):
        return self._n

    def has_edge(self, u, v):
        n = self._n
        return u != v and (self.adj[u * n + v] == "1" or self.adj[v * n + u] == "1")

    def e_cnt(self):
        return sum(self.has_edge(i, j) for i in range(self._n) for j in range(i))

    def automorphisms(self):
        out = C.c_int64()
        check(lib().gat_pattern_automorphisms(self.adj.encode(), C.byref(out)))
        return out.value

    def __repr__(self):
        return "Pattern(%r)" % self.adj


def schedule_candidates(pattern, v_cnt=5000, e_cnt=1000000, tri_cnt=10000, limit=64):
    """Schedule candidates ranked by the cost estimate under the given graph
    statistics (the defaults are the reference's fixed numbers, under which
    entry 0 is Schedule(pattern)). Returns (list of Schedule, total)."""
    adj = Pattern(pattern).adj
    buf = (ScheduleData * max(1, limit))()
    n = C.c_int32()
    check(lib().gat_schedule_candidates(adj.encode(), int(v_cnt), int(e_cnt), int(tri_cnt), buf,
                                        C.sizeof(ScheduleData), int(limit), C.byref(n)))
    return [Schedule.from_data(adj, buf[i]) for i in range(min(limit, n.value))], n.value


class Schedule:
    """Core::Schedule (core/schedule.hpp:592-958), built by libgat.

    Attributes follow the reference's members (prefixs, prefixs_father,
    prefixs_size_only, loop_set_prefix_id, vertex_prefix_start,
    restrictions_start, restrictions, iep_info ...)."""

    def __init__(self, pattern):
        self.pattern = Pattern(pattern)
        self.data = ScheduleData()
        check(lib().gat_schedule_build(self.pattern.adj.encode(), C.byref(self.data), C.sizeof(ScheduleData)))
        self._unpack()

    @classmethod
    def from_data(cls, pattern, data):
        s = cls.__new__(cls)
        s.pattern = Pattern(pattern)
        s.data = data
        s._unpack()
        return s

    def _unpack(self):
        d = self.data
        self.iep_suffix_vertexes = d.iep_suffix_vertexes
        self.basic_vertexes = d.basic_vertexes
        np_ = d.total_prefix_num
        self.prefixs = [list(d.prefix_data[i][:d.prefix_depth[i]]) for i in range(np_)]
        self.prefixs_father = list(d.prefixs_father[:np_])
        self.prefixs_size_only = list(d.prefixs_size_only[:np_])
        b = d.basic_vertexes
        self.loop_set_prefix_id = list(d.loop_set_prefix_id[:b])
        self.vertex_prefix_start = list(d.vertex_prefix_start[:b + 1])
        self.restrictions_start = list(d.restrictions_start[:b + 1])
        nr = d.total_restrictions_num
        self.restrictions = list(zip(d.restrictions[:nr], d.restriction_second[:nr]))
        na = d.n_all_restrictions
        self.all_restrictions = list(zip(d.all_restr_first[:na], d.all_restr_second[:na]))
        self.perm = list(d.perm[:d.v_cnt])
        self.iep_vertex_id = list(d.iep_vertex_id[:d.iep_prefix_num])
        ns = d.subgroups_num
        self.iep_ans_pos = list(d.iep_ans_pos[:ns])
        self.iep_coef = list(d.iep_coef[:ns])
        self.iep_flag = [bool(x) for x in d.iep_flag[:ns]]
        self.automorphisms = d.automorphisms
        self.best_cost = d.best_cost

    def as_dict(self):
        return {k: getattr(self, k) for k in (
            "iep_suffix_vertexes", "basic_vertexes", "prefixs", "prefixs_father", "prefixs_size_only",
            "loop_set_prefix_id", "vertex_prefix_start", "restrictions_start", "restrictions",
            "all_restrictions", "perm", "iep_vertex_id", "iep_ans_pos", "iep_coef", "iep_flag",
            "automorphisms")}


class GlobalMemoryGraph:
    """CSR data graph (infra/graph_backend.cuh:16-170): int64 offsets, int32
    sorted adjacency. Owns a libgat handle."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib._lib is not None:
            lib().gat_graph_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    # ---- constructors ----
    @classmethod
    def _make(cls, fn, *args):
        h = C.c_void_p()
        check(fn(*args, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path, binary=True):
        fn = lib().gat_graph_load_bin if binary else lib().gat_graph_load_txt
        return cls._make(fn, str(path).encode())

    @classmethod
    def erdos_renyi(cls, n, p, seed):
        return cls._make(lib().gat_graph_gen_er, n, p, seed)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0):
        return cls._make(lib().gat_graph_gen_rmat, scale, edge_factor, a, b, c, seed)

    @classmethod
    def chung_lu(cls, n, mean_degree, gamma, max_degree, seed):
        return cls._make(lib().gat_graph_gen_chung_lu, n, mean_degree, gamma, max_degree, seed)

    @classmethod
    def from_edges(cls, n, src, dst):
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        return cls._make(lib().gat_graph_from_edges, n, len(s), s.ctypes.data, d.ctypes.data)

    @classmethod
    def from_csr(cls, offsets, adj):
        o = np.ascontiguousarray(offsets, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int32)
        return cls._make(lib().gat_graph_from_csr, len(o) - 1, len(a), o.ctypes.data,
                         a.ctypes.data if len(a) else None)

    # ---- accessors (graph_backend.cuh:120-138) ----
    def _info(self):
        n, m, d = C.c_int32(), C.c_int64(), C.c_int32()
        check(lib().gat_graph_info(self._h, C.byref(n), C.byref(m), C.byref(d)))
        return n.value, m.value, d.value

    def v_cnt(self):
        return self._info()[0]

    def e_cnt(self):
        """Directed entry count, as the reference stores it (= 2 * edges)."""
        return self._info()[1]

    def undirected_edges(self):
        return self._info()[1] // 2

    def max_degree(self):
        return self._info()[2]

    def csr(self):
        """(offsets, adj) as read-only numpy views into the library's host
        copy; the views keep this graph alive."""
        n, m, _ = self._info()
        po, pa = C.c_void_p(), C.c_void_p()
        check(lib().gat_graph_csr(self._h, C.byref(po), C.byref(pa)))

        def view(ctype, count, addr):
            buf = (ctype * count).from_address(addr)
            buf._graph = self  # the array's base holds the graph, so its memory outlives every view
            a = np.ctypeslib.as_array(buf)
            a.flags.writeable = False
            return a

        off = view(C.c_int64, n + 1, po.value)
        adj = view(C.c_int32, m, pa.value) if m else np.zeros(0, dtype=np.int32)
        return off, adj

    def triangles(self):
        """Number of triangles (host count, cached by the library)."""
        out = C.c_int64()
        check(lib().gat_graph_triangles(self._h, C.byref(out)))
        return out.value

    def degree_histogram(self):
        bins = np.zeros(32, dtype=np.int64)
        check(lib().gat_graph_degree_histogram(self._h, bins.ctypes.data))
        return bins

    def save(self, path):
        check(lib().gat_graph_save_bin(self._h, str(path).encode()))


Graph = GlobalMemoryGraph


SCHEDULES = {"reference": SCHEDULE_REFERENCE, "graph": SCHEDULE_GRAPH, "ranked": SCHEDULE_RANKED}


def _stats_dict(st):
    out = {}
    for k, _ in CountStats._fields_:
        v = getattr(st, k)
        out[k] = list(v) if hasattr(v, "_length_") else v
    return out


class Plan:
    """One or more patterns compiled for one graph (gat_plan).

    schedule="reference" uses the reference's Schedule for every pattern;
    schedule="graph" (unique/labeled modes) ranks the Schedule candidates by
    the reference's cost estimate under this graph's statistics and takes
    the `schedule_rank`-th (an int for every pattern, or a list with one
    rank per pattern); schedule="ranked" ranks them under the reference's
    fixed statistics, so a rank names the same Schedule on every graph."""

    def __init__(self, graph, patterns, mode="unique", schedule="reference", schedule_rank=0):
        if isinstance(patterns, (str, Pattern)):
            patterns = [patterns]
        self.patterns = [Pattern(p).adj for p in patterns]
        self.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        self.graph = graph  # keep alive
        arr = (C.c_char_p * len(self.patterns))(*[p.encode() for p in self.patterns])
        opt = PlanOptions()
        opt.schedule_policy = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
        if isinstance(schedule_rank, (list, tuple)):
            if len(schedule_rank) != len(self.patterns):
                raise ValueError("%d schedule ranks for %d patterns" % (len(schedule_rank), len(self.patterns)))
            self._ranks = (C.c_int32 * len(schedule_rank))(*[int(r) for r in schedule_rank])
            opt.schedule_ranks = self._ranks
        else:
            opt.schedule_rank = int(schedule_rank)
        h = C.c_void_p()
        check(lib().gat_plan_create_ex(graph.handle, arr, len(self.patterns), self.mode, C.byref(opt),
                                       C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib._lib is not None:
            lib().gat_plan_free(h)
            self._h = None

    def schedule(self, i=0):
        d = ScheduleData()
        check(lib().gat_plan_schedule(self._h, i, C.byref(d), C.sizeof(ScheduleData)))
        return Schedule.from_data(self.patterns[i], d)

    def describe(self):
        buf = C.create_string_buffer(1 << 16)
        n = check(lib().gat_plan_describe(self._h, buf, len(buf)))
        if n >= len(buf):
            buf = C.create_string_buffer(n + 1)
            lib().gat_plan_describe(self._h, buf, len(buf))
        return buf.value.decode()

    def count(self, config=None, account=False, graph_replays=0, timeline=False):
        """Counts (list of int, one per pattern) and a stats dict. With
        timeline, the stats hold "timeline": the per-wave / per-workgroup
        records of gat_plan_timeline (an int64 array of shape [n, 8]) and
        "ticks_per_ms"."""
        n = len(self.patterns)
        out = (C.c_uint64 * n)()
        st = CountStats()
        cfg = config if isinstance(config, KernelConfig) else KernelConfig.from_dict(config or {})
        if graph_replays:
            check(lib().gat_count_graph(self._h, C.byref(cfg), int(graph_replays), out, C.byref(st)))
        else:
            flags = (_lib.FLAG_ACCOUNT if account else 0) | (_lib.FLAG_TIMELINE if timeline else 0)
            check(lib().gat_count(self._h, C.byref(cfg), flags, out, C.byref(st)))
        stats = _stats_dict(st)
        if timeline and not graph_replays:
            stats["timeline"], stats["ticks_per_ms"] = self.timeline()
        return [int(x) for x in out], stats

    def timeline(self):
        """Records of the last count run with timeline=True: int64 [n, 8]
        (launch, unit, start, end, build cycles, count cycles, items, pairs)
        and wall-clock ticks per millisecond."""
        import numpy as np
        n, tpm = C.c_int64(), C.c_double()
        check(lib().gat_plan_timeline(self._h, None, 0, C.byref(n), C.byref(tpm)))
        buf = np.zeros((max(1, n.value), 8), dtype=np.int64)
        check(lib().gat_plan_timeline(self._h, buf.ctypes.data, n.value, C.byref(n), C.byref(tpm)))
        return buf[:n.value], tpm.value

    def count_info(self, i=0):
        """{"divisor", "multiplier", "relabeled", "restrictions": [(a, b), ...]}
        of pattern i: the raw count is the number of injective, edge-preserving
        maps psi of the Schedule-ordered pattern with psi(b) < psi(a) for every
        restriction (unique and labeled modes), and count() reports
        raw / divisor * multiplier."""
        d = _lib.CountInfo()
        check(lib().gat_plan_count_info(self._h, i, C.byref(d), C.sizeof(d)))
        return {"divisor": int(d.divisor), "multiplier": int(d.multiplier), "relabeled": bool(d.relabeled),
                "restrictions": [(d.restr_first[k], d.restr_second[k]) for k in range(d.n_restrictions)]}

    def count_shard(self, root_begin, root_end, config=None, account=False, timeline=False):
        """Raw sums (one per pattern) of the embeddings whose Schedule vertex 0
        maps to a root in [root_begin, root_end) of the graph the plan runs on,
        and a stats dict. Over a partition of the roots the raw sums add up to
        divisor * count."""
        n = len(self.patterns)
        out = (C.c_uint64 * n)()
        st = CountStats()
        cfg = config if isinstance(config, KernelConfig) else KernelConfig.from_dict(config or {})
        flags = (_lib.FLAG_ACCOUNT if account else 0) | (_lib.FLAG_TIMELINE if timeline else 0)
        check(lib().gat_count_shard(self._h, C.byref(cfg), flags, int(root_begin), int(root_end), out,
                                    C.byref(st)))
        stats = _stats_dict(st)
        if timeline:
            stats["timeline"], stats["ticks_per_ms"] = self.timeline()
        return [int(x) for x in out], stats


    def count_shard_edges(self, edge_begin, edge_end, config=None, account=False):
        """Raw sums (one per pattern) of the embeddings whose first two
        Schedule vertices map to a directed edge with index in [edge_begin,
        edge_end) of the CSR the plan runs on (array store only), and a
        stats dict. Over a partition of the edges the raw sums add up to
        divisor * count."""
        n = len(self.patterns)
        out = (C.c_uint64 * n)()
        st = CountStats()
        cfg = config if isinstance(config, KernelConfig) else KernelConfig.from_dict(config or {})
        flags = _lib.FLAG_ACCOUNT if account else 0
        check(lib().gat_count_shard_edges(self._h, C.byref(cfg), flags, int(edge_begin), int(edge_end), out,
                                          C.byref(st)))
        return [int(x) for x in out], _stats_dict(st)

    def run_offsets(self):
        """int64 offsets of the CSR the plan runs on: the input's for
        reference-mode plans, else the degree-ordered graph's (prefix sums of
        the degrees sorted in descending order)."""
        import numpy as np
        off, _ = self.graph.csr()
        deg = np.diff(np.asarray(off, dtype=np.int64))
        if self.count_info(0)["relabeled"]:
            deg = np.sort(deg)[::-1]
        return np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)


def count(graph, patterns, mode="unique", config=None, account=False):
    """Exact counts of every pattern in one batched launch."""
    plan = Plan(graph, patterns, mode)
    return plan.count(config, account=account)


class DeviceContext:
    """Engine::DeviceContext{schedule, graph} (engine/context.cuh:10-26)."""

    def __init__(self, schedule, graph_backend):
        self.schedule_data = schedule
        self.graph_backend = graph_backend
        self._plan = None

    def to_device(self):
        # Device buffers are created on the first count; compiling the plan
        # here keeps the reference's call order meaningful.
        if self._plan is None:
            self._plan = Plan(self.graph_backend, [self.schedule_data.pattern.adj], "reference")
        return self


class Executor:
    """Engine::Executor (engine/engine.cuh:124-229)."""

    def __init__(self, set_size, device_type=GPU_DEVICE, config=None):
        if device_type != GPU_DEVICE:
            raise ValueError("only GPU_DEVICE is implemented (the reference's CPU storage path is unused)")
        self.set_size = set_size
        self.config = config
        self.last_stats = None

    def perform_search(self, context):
        context.to_device()
        counts, stats = context._plan.count(self.config)
        self.last_stats = stats
        return counts[0]


def default_config():
    c = KernelConfig()
    check(lib().gat_default_config(C.byref(c)))
    return c


def clique_counts_noalloc(d_offsets, d_adj, v_cnt, kmax, config=None, stream=None):
    """#K3 .. #K_kmax of a CSR graph already in device memory (device
    addresses as ints) through the allocation-free entry point
    gat_count_cliques_noalloc; `stream` is a hipStream_t address or None."""
    n = int(kmax) - 2
    out = (C.c_uint64 * max(1, n))()
    cfg = config if isinstance(config, KernelConfig) else KernelConfig.from_dict(config or {})
    check(lib().gat_count_cliques_noalloc(C.c_void_p(int(d_offsets)), C.c_void_p(int(d_adj)), int(v_cnt), int(kmax),
                                          C.byref(cfg), C.c_void_p(stream) if stream else None, out))
    return [int(x) for x in out[:n]]


def device_count():
    n = C.c_int32()
    lib().gat_device_count(C.byref(n))
    return n.value
