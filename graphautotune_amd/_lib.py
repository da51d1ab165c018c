"""ctypes binding of libgat.so (the C ABI declared in include/gat.h).

This is the reference-side binding a Python caller (for example the tuning
driver) uses: plain pointers and sizes, no torch types. The library must have
been built in-tree (`graphautotune_amd/build.py`); loading fails loudly
otherwise -- there is no CPU fallback for counting.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgat.so")

GAT_OK = 0
GAT_EINVAL = -1
GAT_EIO = -2
GAT_ENOMEM = -3
GAT_ENODEV = -4
GAT_ELIMIT = -5
GAT_EUNSUPPORTED = -6
GAT_EINFEASIBLE = -7
GAT_EINTERNAL = -8

MODE_UNIQUE = 0
MODE_REFERENCE = 1
MODE_LABELED = 2

FLAG_ACCOUNT = 1
FLAG_TIMELINE = 2

MAXV = 10
MAXP = 16
MAXG = 64


class ScheduleData(C.Structure):
    """gat_schedule_data (superset of Core::ScheduleData, schedule.hpp:984-1026)."""
    _fields_ = [
        ("v_cnt", C.c_int32),
        ("iep_suffix_vertexes", C.c_int32),
        ("basic_vertexes", C.c_int32),
        ("total_prefix_num", C.c_int32),
        ("total_restrictions_num", C.c_int32),
        ("prefix_depth", C.c_int32 * MAXP),
        ("prefix_data", (C.c_int32 * MAXV) * MAXP),
        ("prefixs_father", C.c_int32 * MAXP),
        ("prefixs_size_only", C.c_int32 * MAXP),
        ("vertex_prefix_start", C.c_int32 * (MAXV + 1)),
        ("loop_set_prefix_id", C.c_int32 * MAXV),
        ("restrictions_start", C.c_int32 * (MAXV + 1)),
        ("restrictions", C.c_int32 * (MAXV * MAXV)),
        ("restriction_second", C.c_int32 * (MAXV * MAXV)),
        ("iep_prefix_num", C.c_int32),
        ("subgroups_num", C.c_int32),
        ("iep_vertex_id", C.c_int32 * MAXP),
        ("iep_ans_pos", C.c_int32 * MAXG),
        ("iep_coef", C.c_int32 * MAXG),
        ("iep_flag", C.c_int32 * MAXG),
        ("perm", C.c_int32 * MAXV),
        ("n_all_restrictions", C.c_int32),
        ("all_restr_first", C.c_int32 * (MAXV * MAXV)),
        ("all_restr_second", C.c_int32 * (MAXV * MAXV)),
        ("automorphisms", C.c_int32),
        ("adj_rows", C.c_uint32 * MAXV),
        ("best_cost", C.c_double),
    ]


class KernelConfig(C.Structure):
    """gat_kernel_config: runtime form of the reference's compile-time Config."""
    _fields_ = [
        ("block_waves", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("queue_chunk", C.c_int32),
        ("slot_capacity", C.c_int32),
        ("leaf_mode", C.c_int32),
        ("intersect_strategy", C.c_int32),
        ("gallop_ratio", C.c_int32),
        ("search_linear_threshold", C.c_int32),
        ("max_regs", C.c_int32),
        ("vertex_store", C.c_int32),
        ("hub_degree", C.c_int32),
        ("split_us", C.c_int32),
    ]

    KNOBS = ("block_waves", "waves_per_cu", "queue_chunk", "slot_capacity", "leaf_mode",
             "intersect_strategy", "gallop_ratio", "search_linear_threshold", "max_regs", "vertex_store",
             "hub_degree", "split_us")

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.KNOBS}

    @classmethod
    def from_dict(cls, d):
        c = cls()
        for k, v in d.items():
            setattr(c, k, int(v))
        return c


class CountStats(C.Structure):
    _fields_ = [
        ("seconds", C.c_double),
        ("main_kernel_ms", C.c_double),
        ("second_kernel_ms", C.c_double),
        ("overflow_tasks", C.c_uint64),
        ("alg_bytes", C.c_uint64),
        ("alg_bytes_main", C.c_uint64),
        ("alg_bytes_second", C.c_uint64),
        ("tasks", C.c_uint64),
        ("grid_blocks", C.c_int32),
        ("block_threads", C.c_int32),
        ("lds_bytes", C.c_int32),
        ("split_tasks", C.c_int32),
        ("split_overflow", C.c_int32),
        ("split_level", C.c_int64 * 4),
    ]


class PlanOptions(C.Structure):
    _fields_ = [
        ("schedule_policy", C.c_int32),
        ("schedule_rank", C.c_int32),
        ("schedule_ranks", C.POINTER(C.c_int32)),
        ("reserved", C.c_int64 * 2),
    ]


SCHEDULE_REFERENCE, SCHEDULE_GRAPH, SCHEDULE_RANKED = 0, 1, 2


class CountInfo(C.Structure):
    """gat_pattern_count_info: divisor, multiplier and restrictions of one
    pattern of a plan."""
    _fields_ = [
        ("divisor", C.c_uint64),
        ("multiplier", C.c_uint64),
        ("relabeled", C.c_int32),
        ("n_restrictions", C.c_int32),
        ("restr_first", C.c_int32 * (MAXV * MAXV)),
        ("restr_second", C.c_int32 * (MAXV * MAXV)),
    ]


class GatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gat error %d: %s" % (code, msg))
        self.code = code


_P = C.c_void_p
_SIGS = {
    "gat_last_error": (C.c_char_p, []),
    "gat_version": (C.c_char_p, []),
    "gat_schedule_build": (C.c_int, [C.c_char_p, C.POINTER(ScheduleData), C.c_size_t]),
    "gat_schedule_candidates": (C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.c_int64, _P, C.c_size_t,
                                          C.c_int32, C.POINTER(C.c_int32)]),
    "gat_pattern_automorphisms": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    "gat_graph_load_bin": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "gat_graph_load_txt": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "gat_graph_save_bin": (C.c_int, [_P, C.c_char_p]),
    "gat_graph_from_csr": (C.c_int, [C.c_int32, C.c_int64, _P, _P, C.POINTER(_P)]),
    "gat_graph_from_edges": (C.c_int, [C.c_int32, C.c_int64, _P, _P, C.POINTER(_P)]),
    "gat_graph_gen_er": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, C.POINTER(_P)]),
    "gat_graph_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.POINTER(_P)]),
    "gat_graph_gen_chung_lu": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_uint64,
                                         C.POINTER(_P)]),
    "gat_graph_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "gat_graph_csr": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "gat_graph_degree_histogram": (C.c_int, [_P, _P]),
    "gat_graph_triangles": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gat_graph_free": (None, [_P]),
    "gat_plan_create": (C.c_int, [_P, C.POINTER(C.c_char_p), C.c_int32, C.c_int32, C.POINTER(_P)]),
    "gat_plan_create_ex": (C.c_int, [_P, C.POINTER(C.c_char_p), C.c_int32, C.c_int32, C.POINTER(PlanOptions),
                                     C.POINTER(_P)]),
    "gat_plan_schedule": (C.c_int, [_P, C.c_int32, C.POINTER(ScheduleData), C.c_size_t]),
    "gat_count": (C.c_int, [_P, C.POINTER(KernelConfig), C.c_uint32, C.POINTER(C.c_uint64),
                            C.POINTER(CountStats)]),
    "gat_count_graph": (C.c_int, [_P, C.POINTER(KernelConfig), C.c_int32, C.POINTER(C.c_uint64),
                                  C.POINTER(CountStats)]),
    "gat_plan_count_info": (C.c_int, [_P, C.c_int32, C.POINTER(CountInfo), C.c_size_t]),
    "gat_count_shard": (C.c_int, [_P, C.POINTER(KernelConfig), C.c_uint32, C.c_int32, C.c_int32,
                                  C.POINTER(C.c_uint64), C.POINTER(CountStats)]),
    "gat_count_shard_edges": (C.c_int, [_P, C.POINTER(KernelConfig), C.c_uint32, C.c_int64, C.c_int64,
                                        C.POINTER(C.c_uint64), C.POINTER(CountStats)]),
    "gat_count_cliques_noalloc": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.POINTER(KernelConfig), _P,
                                            C.POINTER(C.c_uint64)]),
    "gat_plan_free": (None, [_P]),
    "gat_plan_timeline": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "gat_plan_describe": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "gat_default_config": (C.c_int, [C.POINTER(KernelConfig)]),
    "gat_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "gat_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_P)]),
    "gat_mlp_param_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gat_mlp_get_params": (C.c_int, [_P, _P, C.c_int64]),
    "gat_mlp_set_params": (C.c_int, [_P, _P, C.c_int64]),
    "gat_mlp_forward": (C.c_int, [_P, _P, C.c_int32, _P]),
    "gat_mlp_gradients": (C.c_int, [_P, _P, _P, C.c_int32, _P, C.POINTER(C.c_float)]),
    "gat_mlp_train": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, C.c_float, C.POINTER(C.c_float)]),
    "gat_mlp_free": (None, [_P]),
    "gat_copy_probe": (C.c_int, [C.c_uint64, C.c_int32, C.POINTER(C.c_double)]),
}
# Declared in include/gat.h and exported; tests check every one resolves.
EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """Load libgat.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgat.so not built: run `python -m graphautotune_amd.build` "
                              "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc is not None and rc < 0:
        msg = lib().gat_last_error()
        raise GatError(rc, msg.decode() if msg else "")
    return rc
