"""ctypes wrapper of liboracle.so (checker only; see oracle.h)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build
            build.build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.orc_schedule.argtypes = [C.c_char_p, P]
        L.orc_schedule.restype = C.c_int
        L.orc_automorphisms.argtypes = [C.c_char_p]
        L.orc_automorphisms.restype = C.c_int64
        L.orc_graph_cnt.argtypes = [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_graph_cnt.restype = None
        for f in (L.orc_naive_count, L.orc_labeled_count):
            f.argtypes = [C.c_int32, P, P, C.c_char_p]
            f.restype = C.c_uint64
        L.orc_reference_count.argtypes = [C.c_int32, P, P, P, C.POINTER(C.c_int)]
        L.orc_reference_count.restype = C.c_uint64
        L.orc_reference_count_roots.argtypes = [C.c_int32, P, P, P, C.c_int32, C.c_int32, C.POINTER(C.c_int)]
        L.orc_reference_count_roots.restype = C.c_uint64
        L.orc_clique_walg.argtypes = [C.c_int32, P, P, C.c_int, C.POINTER(C.c_uint64)]
        L.orc_clique_walg.restype = C.c_uint64
        L.orc_clique_walg_sample.argtypes = [C.c_int32, P, P, C.c_int, C.c_int32, C.c_int32,
                                             C.POINTER(C.c_uint64)]
        L.orc_clique_walg_sample.restype = C.c_uint64
        L.orc_clique_bitmap_walg.argtypes = [C.c_int32, P, P, C.c_int, C.POINTER(C.c_uint64)]
        L.orc_clique_bitmap_walg.restype = C.c_uint64
        L.orc_clique_bitmap_walg_sample.argtypes = [C.c_int32, P, P, C.c_int, C.c_int32, C.c_int32,
                                                    C.POINTER(C.c_uint64)]
        L.orc_clique_bitmap_walg_sample.restype = C.c_uint64
        L.orc_restricted_count.argtypes = [C.c_int32, P, P, C.c_int, P, C.c_int, P, P, C.c_int32, C.c_int32]
        L.orc_restricted_count.restype = C.c_uint64
        L.orc_clique_counts_mt.argtypes = [C.c_int32, P, P, C.c_int, C.c_int32, C.c_int32, C.c_int, P]
        L.orc_clique_counts_mt.restype = C.c_int
        L.orc_relabel.argtypes = [C.c_int32, P, P, P, P]
        L.orc_relabel.restype = None
        _lib = L
    return _lib


def _csr(off, adj):
    off = np.ascontiguousarray(off, dtype=np.int64)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    return off, adj, len(off) - 1


def schedule(adj_str, struct_type):
    """Restated Core::Schedule; returns a struct_type (gat_schedule_data layout)."""
    d = struct_type()
    rc = lib().orc_schedule(adj_str.encode(), C.byref(d))
    if rc != 0:
        raise ValueError("oracle schedule failed (%d) for %s" % (rc, adj_str))
    return d


def automorphisms(adj_str):
    return int(lib().orc_automorphisms(adj_str.encode()))


def graph_cnt(k):
    e, o = C.c_int64(), C.c_int64()
    lib().orc_graph_cnt(k, C.byref(e), C.byref(o))
    return e.value, o.value


def naive_count(off, adj, pattern):
    off, adj, n = _csr(off, adj)
    return int(lib().orc_naive_count(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, pattern.encode()))


def labeled_count(off, adj, pattern):
    off, adj, n = _csr(off, adj)
    return int(lib().orc_labeled_count(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, pattern.encode()))


def unique_count(off, adj, pattern):
    lab = labeled_count(off, adj, pattern)
    aut = automorphisms(pattern)
    assert lab % aut == 0, (lab, aut)
    return lab // aut


def reference_count(off, adj, sched_struct):
    """Reference engine value (long long bits as uint64) and a defined flag."""
    off, adj, n = _csr(off, adj)
    ok = C.c_int()
    v = lib().orc_reference_count(n, off.ctypes.data, adj.ctypes.data if len(adj) else None,
                                  C.byref(sched_struct), C.byref(ok))
    return int(v), bool(ok.value)


def restricted_count(off, adj, rows, restrictions, root_lo=0, root_hi=None):
    """Injective, edge-preserving maps psi of the pattern whose vertex i is
    adjacent to j iff bit j of rows[i], with psi(b) < psi(a) for every (a, b)
    in restrictions, over the roots psi(0) in [root_lo, root_hi)."""
    off, adj, n = _csr(off, adj)
    k = len(rows)
    r = (C.c_uint32 * max(1, k))(*rows)
    nr = len(restrictions)
    fa = (C.c_int32 * max(1, nr))(*[a for a, _ in restrictions])
    sb = (C.c_int32 * max(1, nr))(*[b for _, b in restrictions])
    hi = n if root_hi is None else root_hi
    return int(lib().orc_restricted_count(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, k, r, nr,
                                          fa, sb, int(root_lo), int(hi)))


def reference_count_roots(off, adj, sched_struct, root_lo, root_hi):
    """reference_count over first vertices v0 in [root_lo, root_hi)."""
    off, adj, n = _csr(off, adj)
    ok = C.c_int()
    v = lib().orc_reference_count_roots(n, off.ctypes.data, adj.ctypes.data if len(adj) else None,
                                        C.byref(sched_struct), int(root_lo), int(root_hi), C.byref(ok))
    return int(v), bool(ok.value)


def relabel(off, adj):
    off, adj, n = _csr(off, adj)
    o2 = np.zeros(n + 1, dtype=np.int64)
    a2 = np.zeros(max(1, len(adj)), dtype=np.int32)
    lib().orc_relabel(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, o2.ctypes.data, a2.ctypes.data)
    return o2, a2[:len(adj)]


def clique_walg(off_relabeled, adj_relabeled, k):
    off, adj, n = _csr(off_relabeled, adj_relabeled)
    w = C.c_uint64()
    c = lib().orc_clique_walg(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, k, C.byref(w))
    return int(c), int(w.value)


def clique_walg_sample(off_relabeled, adj_relabeled, k, stride, first=0):
    off, adj, n = _csr(off_relabeled, adj_relabeled)
    w = C.c_uint64()
    c = lib().orc_clique_walg_sample(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, k, stride, first,
                                     C.byref(w))
    return int(c), int(w.value)


def clique_bitmap_walg(off_relabeled, adj_relabeled, kmax):
    """(#K_kmax, W_alg of the bitmap path for cliques K3..kmax)."""
    off, adj, n = _csr(off_relabeled, adj_relabeled)
    w = C.c_uint64()
    c = lib().orc_clique_bitmap_walg(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, kmax, C.byref(w))
    return int(c), int(w.value)


def clique_bitmap_walg_sample(off_relabeled, adj_relabeled, kmax, stride, first=0):
    """(#K_kmax with top vertex in first, first + stride, ..., W_alg of the
    bitmap store over those top vertices)."""
    off, adj, n = _csr(off_relabeled, adj_relabeled)
    w = C.c_uint64()
    c = lib().orc_clique_bitmap_walg_sample(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, kmax,
                                            stride, first, C.byref(w))
    return int(c), int(w.value)


def clique_counts_mt(off, adj, kmax, threads, first=0, stride=1):
    """[#K3, ..., #K_kmax] rooted at top vertices first, first + stride, ...
    (the whole graph by default), counted on `threads` host threads."""
    off, adj, n = _csr(off, adj)
    out = (C.c_uint64 * 7)()
    rc = lib().orc_clique_counts_mt(n, off.ctypes.data, adj.ctypes.data if len(adj) else None, kmax, first, stride,
                                    threads, out)
    if rc != 0:
        raise ValueError("orc_clique_counts_mt failed (%d)" % rc)
    return [int(out[k]) for k in range(3, kmax + 1)]
