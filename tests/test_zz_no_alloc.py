"""The allocation-free clique count, gat_count_cliques_noalloc (GPU, through
the C ABI): on a graph uploaded by the test, its counts equal gat_count's
(the bitmap clique store of a plan) on two calls in a row, the second on a
smaller graph, and an input beyond its limits is refused with GAT_ELIMIT
while the next call still counts right (the scratch array is cleared per
call)."""
import ctypes as C
import os

import numpy as np
import pytest

from patterns import clique

pytestmark = pytest.mark.gpu


class _Device:
    """Device buffers through the HIP runtime libgat.so is linked against
    (already loaded by the gat fixture, so this is the same runtime)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | C.RTLD_GLOBAL)
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(arr.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _say(*a):
    print(*a, flush=True)


def test_noalloc_matches_gat_count(gat, orc, gpu):
    dev = _Device()
    try:
        # two calls in a row, the second on a smaller input
        for scale, kmax in ((16, 5), (11, 6)):
            g = gat.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, 20)
            off, adj = g.csr()
            ro, ra = orc.relabel(off, adj)  # degree-descending ids, as the plan's own graph
            pats = [clique(k) for k in range(3, kmax + 1)]
            want, _ = gat.Plan(g, pats, "unique").count({"vertex_store": 2})
            d_off, d_adj = dev.upload(ro.astype(np.int64)), dev.upload(ra.astype(np.int32))
            _say("noalloc", "rmat%d" % scale, "kmax", kmax)
            got = gat.core.clique_counts_noalloc(d_off, d_adj, g.v_cnt(), kmax)
            _say("noalloc", "rmat%d" % scale, got, want)
            assert got == want, (scale, got, want)
            # a second config on the same scratch: same counts
            got2 = gat.core.clique_counts_noalloc(d_off, d_adj, g.v_cnt(), kmax,
                                                  {"slot_capacity": 64, "queue_chunk": 4, "block_waves": 2})
            assert got2 == want, (scale, got2, want)
    finally:
        dev.free()


def test_noalloc_limits(gat, orc, gpu):
    from graphautotune_amd._lib import GAT_ELIMIT, GatError
    dev = _Device()
    try:
        # K_1100: the highest vertex has 1,099 neighbours below it (> 1,024)
        m = 1100
        off = np.arange(m + 1, dtype=np.int64) * (m - 1)
        adj = np.array([u for v in range(m) for u in range(m) if u != v], dtype=np.int32)
        d_off, d_adj = dev.upload(off), dev.upload(adj)
        with pytest.raises(GatError) as e:
            gat.core.clique_counts_noalloc(d_off, d_adj, m, 4)
        assert e.value.code == GAT_ELIMIT
        with pytest.raises(GatError) as e:
            gat.core.clique_counts_noalloc(d_off, d_adj, (1 << 22) + 1, 4)  # refused before any launch
        assert e.value.code == GAT_ELIMIT
        # the next call on the shared scratch counts right: K_1000 has C(1000, k) k-cliques
        m = 1000
        off = np.arange(m + 1, dtype=np.int64) * (m - 1)
        adj = np.array([u for v in range(m) for u in range(m) if u != v], dtype=np.int32)
        d_off, d_adj = dev.upload(off), dev.upload(adj)
        from math import comb
        got = gat.core.clique_counts_noalloc(d_off, d_adj, m, 4)
        assert got == [comb(m, 3), comb(m, 4)], got
    finally:
        dev.free()
