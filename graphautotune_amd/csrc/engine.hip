// MI355X (gfx950) subgraph pattern-matching engine.
//
// One persistent launch runs every pattern of a plan. Each wavefront (64
// lanes) takes a chunk of directed edges from a global work queue; an edge
// (v0 -> v1) fixes the first two pattern vertices, and the wave walks the
// plan's trie depth-first from there:
//   * non-leaf levels take one candidate at a time; their prefix sets are
//     intersections computed by all 64 lanes (stream one sorted list,
//     binary-search the other, compact with ballot + mbcnt) into LDS;
//   * leaf levels count every candidate at once: the (candidate, element)
//     pairs of all candidates form one flattened index space that the lanes
//     sweep 64 at a time, finding their owner candidate with ds_bpermute and
//     crediting hits with ballot + popcount, so lane utilisation does not
//     depend on set sizes. The IEP tail is then combined per candidate.
// Counts are exact uint64 (modular arithmetic for the signed IEP terms).
// A task whose intersection could outgrow its LDS buffer is abandoned and
// queued for a second launch that keeps the same buffers in HBM.
//
// Replaces the six-kernel host-recursive executor of the reference
// (include/engine/engine.cuh:232-428, include/engine/engine_kernel.cuh).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "handles.hpp"
#include "plan_dev.h"

namespace gatk {

constexpr int W = 64;
constexpr int kSplitLevels = 3;  // LDS subtask levels (see KArgs)
constexpr int kLevelsHbm = kSplitLevels + 1;  // timeline id of the HBM launch (its subtasks: + 1)

// A split subtask: candidates [lo, hi) of the loop of node `node` (at depth
// `depth` >= 2) below root edge `edge` (directed edge index: v0 = its row,
// v1 = its entry), with the candidates of depths 2 .. depth-1 taken from
// loop positions pos[2 ..]. lo >= hi marks an empty or cancelled entry.
struct SplitTask {
    long long edge;
    int node, depth, lo, hi;
    int pos[GATP_MAX_DEPTH];
};

struct KArgs {
    const long long *off;
    const int *adj;
    const int *split;  // split[v] = number of neighbours of v below v
    int n;
    long long m;
    long long e_lo, e_hi;        // root edges of this count: [e_lo, e_hi)
    const GatPlanDev *plan;
    unsigned long long *counts;  // [n_patterns]
    unsigned long long *queue;   // [0] main queue head, [1] overflow head
    long long *ovf;              // overflow task starts
    unsigned long long *ovf_n;   // overflow task count
    long long ovf_cap;
    int *gstore;                 // HBM set buffers (overflow launch)
    unsigned long long *bytes;   // W_alg accumulator
    unsigned int *err;           // device error flags
    int cap;                     // entries per set buffer
    int chunk;
    int lin_thr;
    int gallop_ratio;
    int strategy;
    int account;
    // Dense neighbourhood bitmaps of high-degree vertices ("hubs"): row h
    // holds bit x set iff x is a neighbour of hub h. hub_map == nullptr means
    // the graph is degree-ordered and the hubs are vertices [0, n_hubs).
    const unsigned *hub_bits;
    const int *hub_map;
    long long hub_words;  // words per row: (n + 31) / 32
    int n_hubs;
    // Splitting: a task that has run longer than split_ticks of the wall
    // clock while the queue it came from is drained hands the rest of its
    // shallowest unfinished loop at depth >= 2 to subs[out_base ..] (count in
    // *out_n, at most out_cap used). The main launch makes level-1 subtasks;
    // the launch of level k (sub_mode k, LDS buffers) runs them and may split
    // into level k + 1, up to level kSplitLevels, which does not split.
    // Subtasks that outgrow LDS are copied to the HBM list
    // (subs[sovf_base ..], count *sovf_n), which also takes the splits of the
    // HBM launch; a last subtask launch with HBM buffers (sub_mode
    // kSplitLevels + 1) runs it.
    int sub_mode;
    SplitTask *subs;
    unsigned long long *in_n, *out_n, *sovf_n;
    long long in_base, out_base, sovf_base;
    long long in_cap, out_cap, sovf_cap;
    long long split_ticks;  // <= 0: no splitting
    // Timeline (GAT_FLAG_TIMELINE): one record of kTlFields int64 per wave
    // at tl + kTlFields * (tl_base + wave), or none when tl is null.
    long long *tl;
    int tl_base;
};

// Timeline record: launch id, unit (wave or workgroup), wall-clock start and
// end, shader cycles spent building rows and counting (clique store), items
// taken from the queue, and pairs counted (clique store).
constexpr int kTlFields = 8;
__device__ __forceinline__ void tl_write(long long *tl, int idx, int launch, int unit, long long t0, long long t1,
                                         long long build, long long count, long long items, long long pairs) {
    long long *r = tl + (long long)kTlFields * idx;
    r[0] = launch;
    r[1] = unit;
    r[2] = t0;
    r[3] = t1;
    r[4] = build;
    r[5] = count;
    r[6] = items;
    r[7] = pairs;
}

// Subtask blocks a task may hand on between commits. A batched plan's trie
// starts a loop for every child node of a candidate, and once a task is due
// to split each such loop goes to the next split launch with one record.
constexpr int kMaxDonations = 64;

struct WaveState {
    unsigned long long task_acc[GATP_MAX_TERMS];
    unsigned long long wave_acc[GATP_MAX_TERMS];
    long long sbeg[GATP_MAX_SLOTS];
    int slen[GATP_MAX_SLOTS];
    int staged[GATP_MAX_SLOTS];  // neighbour list copied into its LDS buffer
    int val[GATP_MAX_DEPTH + 1];
    int f_node[GATP_MAX_DEPTH + 1];
    int f_cur[GATP_MAX_DEPTH + 1];
    int f_hi[GATP_MAX_DEPTH + 1];
    int leafcnt[GATP_MAX_LEAF_OPS][W];
    unsigned long long bytes;   // committed W_alg of this wave
    unsigned long long ebytes;  // W_alg of the work not yet committed
    long long t0;               // wall clock at the start of the task
    long long rb;               // first edge of the task's root v0
    long long sub_edge;         // root edge of the running subtask, or -1
    long long don_base[kMaxDonations];  // subtask blocks donated since the last commit
    int don_cnt[kMaxDonations];
    int n_don;
    int base_depth;             // frames exist from this depth on (2, or a subtask's depth)
    int path_pos[GATP_MAX_DEPTH + 1];  // a subtask's positions at depths 2 .. base_depth-1
    int abort;
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int mbcnt(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long long uni64(long long x) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(x & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(x >> 32));
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ int bperm(int x, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, x); }
__device__ __forceinline__ long long bperm64(long long x, int src) {
    const int lo = bperm((int)(x & 0xffffffffll), src), hi = bperm((int)(x >> 32), src);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// Reads of one lane's registers by the whole wave: LanePull for a lane index
// that is the same in every lane (v_readlane, no LDS crossbar), PermPull for
// a per-lane index (ds_bpermute).
struct LanePull {
    int s;
    __device__ __forceinline__ int operator()(int x) const { return __builtin_amdgcn_readlane(x, s); }
    __device__ __forceinline__ long long operator()(long long x) const {
        const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), s);
        const int hi = __builtin_amdgcn_readlane((int)(x >> 32), s);
        return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
};
struct PermPull {
    int s;
    __device__ __forceinline__ int operator()(int x) const { return bperm(x, s); }
    __device__ __forceinline__ long long operator()(long long x) const { return bperm64(x, s); }
};
// Lanes own consecutive segments [off_l, off_l + len_l) of T elements, off_l
// not decreasing with l; lane j of a chunk of W elements starting at k0
// handles element k0 + j, whose owner is the last lane whose segment starts
// at or before it. The lanes with off_l <= x are a prefix of the wave, so
// one ballot gives the owner of a position that is the same in every lane:
// when the chunk's first and last elements have one owner, so do all its
// elements, and body(LanePull) reads that lane's fields with readlane;
// otherwise each lane finds its owner by binary search over the lanes and
// body(PermPull) reads them with ds_bpermute. Every lane must be active.
template <class F>
__device__ __forceinline__ void by_owner(int off, int k0, int T, F &&body) {
    const int s0 = __popcll(ballot(off <= k0)) - 1;
    const int s1 = __popcll(ballot(off <= min(k0 + W, T) - 1)) - 1;
    if (s0 == s1) {
        body(LanePull{s0});
    } else {
        const int j = k0 + lane_id();
        int s = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
            if (bperm(off, s + step) <= j && s + step < W) s += step;
        body(PermPull{s});
    }
}
// LDS accesses of one wavefront are performed in program order; the fences
// keep the compiler from moving them across this point.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int excl_scan(int x, int *total) {
    int v = x;
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        const int y = __shfl_up(v, d);
        if (lane_id() >= d) v += y;
    }
    *total = __shfl(v, W - 1);
    return v - x;
}
__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
    return x;
}
__device__ __forceinline__ unsigned long long range_mask(int lo, int hi) {  // bits [lo,hi), 0<=lo<hi<=64
    const unsigned long long h = hi >= 64 ? ~0ull : ((1ull << hi) - 1);
    return h & ~((1ull << lo) - 1);
}

// Number of elements of sorted a[0,len) below t.
template <class P>
__device__ __forceinline__ int lbound(P a, int len, int t) {
    int lo = 0, n = len;
    while (n > 0) {
        const int h = n >> 1;
        if (a[lo + h] < t) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}
// Membership of x in sorted a[0,len): linear scan below lin_thr entries
// (the reference's set_search_type switch, array_vertex_set.cuh:145-158),
// otherwise a branch-free binary descent that never reads past len.
template <class P>
__device__ __forceinline__ bool has(P a, int len, int x, int lin_thr) {
    if (len <= 0) return false;
    if (len < lin_thr) {
        for (int i = 0; i < len; i++) {
            const int y = a[i];
            if (y >= x) return y == x;
        }
        return false;
    }
    int base = 0, n = len;
    while (n > 1) {
        const int half = n >> 1;
        base = (a[base + half] <= x) ? base + half : base;
        n -= half;
    }
    return a[base] == x;
}

struct Ctx {
    const KArgs &a;
    const GatPlanDev &P;
    WaveState *S;
    int *store;
    int ln;
};

// Truncation bound of an op/vid: max over groups of min over member levels.
// Level `d` reads the per-lane candidate c.
__device__ __forceinline__ int group_bound(const WaveState *S, const uint8_t *groups, int ngroups, int d, int c) {
    if (ngroups == 0) return INT_MAX;
    int best = INT_MIN;
    for (int g = 0; g < ngroups; g++) {
        int mn = INT_MAX;
        const unsigned m = groups[g];
        for (int f = 0; f <= GATP_MAX_DEPTH; f++)
            if (m >> f & 1u) mn = min(mn, f == d ? c : S->val[f]);
        best = max(best, mn);
    }
    return best;
}
__device__ __forceinline__ int mask_bound(const WaveState *S, unsigned mask, int d, int c) {
    int mn = INT_MAX;
    for (int f = 0; f <= GATP_MAX_DEPTH; f++)
        if (mask >> f & 1u) mn = min(mn, f == d ? c : S->val[f]);
    return mn;
}
__device__ __forceinline__ bool equals_any(const WaveState *S, unsigned mask, int d, int x) {
    bool r = false;
    for (int f = 0; f < d; f++)
        if (mask >> f & 1u) r |= (S->val[f] == x);
    return r;
}

// Length of N(v) below t (split[] answers the common t == v case).
__device__ __forceinline__ int trunc_len(const KArgs &a, int v, long long beg, int deg, int t) {
    if (t == INT_MAX) return deg;
    if (t == v) return a.split[v];
    return lbound(a.adj + beg, deg, t);
}

// Hub row of vertex v, or -1.
__device__ __forceinline__ int hub_of(const KArgs &a, int v) {
    if (a.n_hubs == 0) return -1;
    if (a.hub_map) return a.hub_map[v];
    return v < a.n_hubs ? v : -1;
}
// Membership of x in the first len entries of hub h's sorted neighbour
// list, whose last kept entry is lim: one bitmap word instead of a search.
__device__ __forceinline__ bool hub_has(const KArgs &a, int h, int x, int lim) {
    return x <= lim && ((a.hub_bits[(long long)h * a.hub_words + (x >> 5)] >> (x & 31)) & 1u);
}
// Largest kept entry of a neighbour-list prefix of length len (-1 if empty).
__device__ __forceinline__ int list_lim(const KArgs &a, long long beg, int len) {
    return len > 0 ? a.adj[beg + len - 1] : -1;
}

// Number of distinct values among ex[0,nex) that lie in the first len
// entries of N(v) (list at beg): the excluded chosen vertices a COUNT
// without a father set must subtract from len. One membership test each
// (bitmap word for hubs) instead of a sweep of the list.
__device__ __forceinline__ int excluded_in_list(const KArgs &a, int v, long long beg, int len, const int *ex,
                                                int nex) {
    if (len <= 0) return 0;
    const int h = hub_of(a, v);
    const int lim = h >= 0 ? list_lim(a, beg, len) : 0;
    int k = 0;
    for (int q = 0; q < nex; q++) {
        bool seen = false;
        for (int q2 = 0; q2 < q; q2++) seen |= ex[q2] == ex[q];
        if (seen) continue;
        if (h >= 0 ? hub_has(a, h, ex[q], lim) : has(a.adj + beg, len, ex[q], a.lin_thr)) k++;
    }
    return k;
}

// Wave-cooperative intersection of stream[0,slen) with the set tested by
// probe(x), skipping the (at most 8) excluded values; writes the sorted
// result to out when MAT.
template <bool MAT, class SP, class PR, class OP>
__device__ __forceinline__ int wave_isect(SP stream, int slen, PR probe, OP out, const int *ex, int nex, int ln) {
    int total = 0;
    for (int base = 0; base < slen; base += W) {
        const int i = base + ln;
        bool hit = false;
        int x = 0;
        if (i < slen) {
            x = stream[i];
            hit = probe(x);
            for (int k = 0; k < nex; k++) hit &= (x != ex[k]);
        }
        const unsigned long long m = ballot(hit);
        if (MAT && hit) out[total + mbcnt(m)] = x;
        total += __popcll(m);
    }
    return total;
}

// Merge-path intersection of two sorted lists A[0,na) and B[0,nb), for
// pairs of similar length. In merge order A[i] precedes B[j] iff
// A[i] <= B[j]; the 64 lanes split the na + nb merged positions into equal
// runs, find where their run starts by a binary search along its diagonal,
// and merge their run serially. A common value x = A[i] = B[j] is found
// once, when B[j] is consumed and the A element before it (A[i-1], possibly
// consumed by the previous lane) equals it. Sorted output (MAT): one pass
// counts, an exclusive scan of the lanes' counts places each run's hits,
// and a second pass writes them.
template <bool MAT, class PA, class PB, class OP>
__device__ __forceinline__ int wave_merge_isect(PA A, int na, PB B, int nb, OP out, const int *ex, int nex, int ln) {
    const int total = na + nb;
    if (na == 0 || nb == 0) return 0;
    const int run = (total + W - 1) / W;
    const int k0 = min(total, ln * run), k1 = min(total, k0 + run);
    int lo = max(0, k0 - nb), hi = min(k0, na);
    while (lo < hi) {  // i = number of A elements among the first k0 merged
        const int mid = (lo + hi) >> 1;
        if (A[mid] <= B[k0 - 1 - mid]) lo = mid + 1; else hi = mid;
    }
    const int i0 = lo, j0 = k0 - lo;
    int cnt = 0;
    for (int pass = 0; pass < (MAT ? 2 : 1); pass++) {
        int i = i0, j = j0, w = 0;
        if (pass == 1) {
            int tot = 0;
            w = excl_scan(cnt, &tot);
        }
        int prev = i > 0 ? A[i - 1] : INT_MIN;  // the A element merged last
        int a = i < na ? A[i] : INT_MAX;
        for (int k = k0; k < k1; k++) {
            if (i < na && (j >= nb || a <= B[j])) {
                prev = a;
                i++;
                a = i < na ? A[i] : INT_MAX;
            } else {
                const int x = B[j++];
                bool hit = prev == x;
                for (int q = 0; q < nex; q++) hit &= (x != ex[q]);
                if (hit) {
                    if (pass == 0) cnt++;
                    else out[w++] = x;
                }
            }
        }
    }
    int tot = 0;
    (void)excl_scan(cnt, &tot);
    return tot;
}

// Galloping intersection of a short sorted list S[0,ns) with a long one
// L[0,nl), for skewed pairs. Each lane takes a block of up to 64
// consecutive elements of S (rounds of 64 blocks), finds the first by a
// binary search of L and every further one by an exponential search from
// the previous position (steps 1, 2, 4, ... then a binary search of the
// last step), so an element costs about 2 log2 of its gap in L instead of
// log2(nl). Hits are kept as a per-lane bit mask; the exclusive scan of
// their counts places the sorted output (MAT).
template <bool MAT, class PS, class PL, class OP>
__device__ __forceinline__ int wave_gallop_isect(PS S, int ns, PL L, int nl, OP out, const int *ex, int nex, int ln) {
    int total = 0;
    for (int base = 0; base < ns; base += W * W) {
        const int blk = min(W, (ns - base + W - 1) / W);
        const int b0 = min(ns, base + ln * blk), b1 = min(ns, b0 + blk);
        unsigned long long mask = 0;
        int pos = 0;
        for (int e = b0; e < b1; e++) {
            const int x = S[e];
            int lo, hi;  // the first element >= x lies in [lo, hi]
            if (e == b0) {
                lo = 0;
                hi = nl;
            } else {
                lo = pos;
                int step = 1;
                hi = pos;
                while (hi < nl && L[hi] < x) {
                    lo = hi + 1;
                    hi = pos + step;
                    step <<= 1;
                }
                hi = min(hi, nl);
            }
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (L[mid] < x) lo = mid + 1; else hi = mid;
            }
            pos = lo;
            bool hit = pos < nl && L[pos] == x;
            for (int q = 0; q < nex; q++) hit &= (x != ex[q]);
            if (hit) mask |= 1ull << (e - b0);
        }
        int tot = 0;
        const int w = excl_scan(__popcll(mask), &tot);
        if (MAT) {
            int o = total + w;
            while (mask) {
                const int b = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                out[o++] = S[b0 + b];
            }
        }
        total += tot;
    }
    return total;
}

// Stream A (the new vertex's neighbour list) and probe the father set, or
// the reverse? A side is cheap to probe when it sits in LDS or has a hub
// bitmap; the other side is then streamed unless it is gallop_ratio times
// longer.
__device__ __forceinline__ bool stream_a(const KArgs &a, int alen, int blen, bool b_fast, bool a_fast) {
    if (a.strategy == 1) return true;
    if (a.strategy == 2) return alen <= blen;
    const long long r = a.gallop_ratio;
    if (b_fast && !a_fast) return !((long long)alen > r * (long long)blen);
    if (a_fast && !b_fast) return (long long)blen > r * (long long)alen;
    return alen <= blen;
}

// Calls f with the slot's content pointer; the two branches keep the LDS
// and global provenance apart so each access compiles to ds_* or global_*.
template <bool GS, class F>
__device__ __forceinline__ auto with_slot(const Ctx &C, int slot, F f) {
    const GatSlot &s = C.P.slots[slot];
    if (s.kind == GATP_ALIAS && (GS || !C.S->staged[slot])) return f(C.a.adj + C.S->sbeg[slot]);
    return f(C.store + (long long)s.store * C.a.cap);
}
// Is the slot's content in this wave's LDS buffers?
template <bool GS>
__device__ __forceinline__ bool in_lds(const Ctx &C, int slot) {
    if (GS) return false;
    const GatSlot &s = C.P.slots[slot];
    return s.kind == GATP_LIST || C.S->staged[slot];
}

// Evaluate one op for the wave-uniform candidate c at depth d.
template <bool GS>
__device__ __forceinline__ void exec_op_uniform(const Ctx &C, const GatOp &o, int d, int c, bool dup) {
    const KArgs &a = C.a;
    WaveState *S = C.S;
    const int t = uni(group_bound(S, o.groups, o.ngroups, d, c));
    const long long ab = a.off[c];
    const int deg = (int)(a.off[c + 1] - ab);
    int alen = dup ? 0 : uni(trunc_len(a, c, ab, deg, t));
    const int slot = o.slot;
    if (o.kind == GATP_ALIAS) {
        // a list that later ops search is staged into LDS when it fits
        const bool stage = !GS && o.store >= 0 && alen <= a.cap;
        if (stage) {
            int *dst = C.store + (long long)o.store * a.cap;
            for (int k = C.ln; k < alen; k += W) dst[k] = a.adj[ab + k];
        }
        if (C.ln == 0) {
            S->sbeg[slot] = ab;
            S->slen[slot] = alen;
            S->staged[slot] = stage ? 1 : 0;
        }
        return;
    }
    int ex[8];
    int nex = 0;
    if (o.kind == GATP_COUNT)
        for (int f = 0; f < d; f++)
            if (o.excl >> f & 1u) ex[nex++] = S->val[f];
    const int f = o.father;
    int blen = 0, result = 0;
    const bool has_r = f >= 0;
    bool b_lds = false;
    int hub_b = -1;
    long long bbeg = 0;
    if (has_r) {
        const GatSlot &fs = C.P.slots[f];
        b_lds = in_lds<GS>(C, f);
        const int flen = S->slen[f];
        blen = dup ? 0 : (t == INT_MAX ? flen : uni(with_slot<GS>(C, f, [&](auto B) { return lbound(B, flen, t); })));
        if (fs.kind == GATP_ALIAS) {
            bbeg = S->sbeg[f];
            if (!b_lds) hub_b = uni(hub_of(a, S->val[fs.depth]));
        }
    }
    if (o.kind == GATP_LIST) {
        const int ub = has_r ? min(alen, blen) : alen;
        if (ub > a.cap) {
            if (C.ln == 0) S->abort = 1;
            return;
        }
    }
    const int *A = a.adj + ab;
    int *out = C.store + (long long)(o.store < 0 ? 0 : o.store) * a.cap;
    const int hub_a = has_r ? uni(hub_of(a, c)) : -1;
    const bool sa = !has_r || stream_a(a, alen, blen, b_lds || hub_b >= 0, hub_a >= 0);
    const int lin = a.lin_thr;
    auto run = [&](auto stream, int slen, auto probe) {
        return o.kind == GATP_LIST ? wave_isect<true>(stream, slen, probe, out, ex, nex, C.ln)
                                   : wave_isect<false>(stream, slen, probe, out, ex, nex, C.ln);
    };
    // merge path or galloping between two sorted arrays (strategies 3 and 4;
    // a hub's bitmap is not used there)
    const int sorted_pair = has_r && (a.strategy == 3 || a.strategy == 4) ? a.strategy : 0;
    auto run2 = [&](auto P, int np_, auto Q, int nq) {  // P, Q: sorted arrays
        if (sorted_pair == 3)
            return o.kind == GATP_LIST ? wave_merge_isect<true>(P, np_, Q, nq, out, ex, nex, C.ln)
                                       : wave_merge_isect<false>(P, np_, Q, nq, out, ex, nex, C.ln);
        // galloping: the shorter list is walked, the longer one searched
        if (np_ <= nq)
            return o.kind == GATP_LIST ? wave_gallop_isect<true>(P, np_, Q, nq, out, ex, nex, C.ln)
                                       : wave_gallop_isect<false>(P, np_, Q, nq, out, ex, nex, C.ln);
        return o.kind == GATP_LIST ? wave_gallop_isect<true>(Q, nq, P, np_, out, ex, nex, C.ln)
                                   : wave_gallop_isect<false>(Q, nq, P, np_, out, ex, nex, C.ln);
    };
    if (sorted_pair) {
        if (b_lds) {
            const int *B = C.store + (long long)C.P.slots[f].store * a.cap;
            result = run2(A, alen, B, blen);
        } else {
            const int *B = C.P.slots[f].kind == GATP_ALIAS ? a.adj + bbeg
                                                           : C.store + (long long)C.P.slots[f].store * a.cap;
            result = run2(A, alen, B, blen);
        }
    } else if (!has_r && o.kind == GATP_COUNT) {
        result = alen - uni(excluded_in_list(a, c, ab, alen, ex, nex));
    } else if (!has_r) {
        result = run(A, alen, [](int) { return true; });
    } else if (sa) {
        // stream A, probe the father: LDS buffer, hub bitmap, or HBM list
        if (b_lds) {
            const int *B = C.store + (long long)C.P.slots[f].store * a.cap;
            result = run(A, alen, [&](int x) { return has(B, blen, x, lin); });
        } else if (hub_b >= 0) {
            const int lim = uni(list_lim(a, bbeg, blen));
            result = run(A, alen, [&](int x) { return hub_has(a, hub_b, x, lim); });
        } else {
            const int *B = C.P.slots[f].kind == GATP_ALIAS ? a.adj + bbeg
                                                           : C.store + (long long)C.P.slots[f].store * a.cap;
            result = run(A, alen, [&](int x) { return has(B, blen, x, lin); });
        }
    } else {
        // stream the father, probe A
        auto probe_a = [&](int x) { return has(A, alen, x, lin); };
        const int lim_a = hub_a >= 0 ? uni(list_lim(a, ab, alen)) : 0;
        auto probe_hub = [&](int x) { return hub_has(a, hub_a, x, lim_a); };
        if (b_lds) {
            const int *B = C.store + (long long)C.P.slots[f].store * a.cap;
            result = hub_a >= 0 ? run(B, blen, probe_hub) : run(B, blen, probe_a);
        } else {
            const int *B = C.P.slots[f].kind == GATP_ALIAS ? a.adj + bbeg
                                                           : C.store + (long long)C.P.slots[f].store * a.cap;
            result = hub_a >= 0 ? run(B, blen, probe_hub) : run(B, blen, probe_a);
        }
    }
    if (C.ln == 0) {
        S->slen[slot] = result;
        if (a.account) {
            const GatSlot *fs = has_r ? &C.P.slots[f] : nullptr;
            S->ebytes += 4ull * (alen + blen) + (o.kind == GATP_LIST ? 4ull * result : 0ull) + 16ull +
                        ((fs && fs->kind == GATP_ALIAS) ? 16ull : 0ull);
        }
    }
}

// Subtask blocks handed on since the last commit are emptied when the work
// they came from is abandoned: the HBM launch re-runs that work whole.
__device__ __noinline__ void cancel_donations(const KArgs &a, WaveState *S, int ln) {
    const int nd = S->n_don;
    for (int i = 0; i < nd; i++) {
        const long long b = S->don_base[i];
        const int cnt = S->don_cnt[i];
        for (int k = ln; k < cnt; k += W) {
            a.subs[b + k].lo = 0;
            a.subs[b + k].hi = 0;
        }
    }
    wsync();
    if (ln == 0) S->n_don = 0;
    wsync();
}

// Split only in the tail: the queue the task came from has nothing left for
// idle waves (queue_drained) and the task has outrun its budget (want_split).
template <bool GS>
__device__ __forceinline__ bool queue_drained(const KArgs &a) {
    if (a.split_ticks <= 0 || a.sub_mode >= kSplitLevels) return false;
    if (GS) {  // the HBM launch: its overflow list is drained
        const unsigned long long made = __atomic_load_n(a.ovf_n, __ATOMIC_RELAXED);
        const unsigned long long total = made < (unsigned long long)a.ovf_cap ? made : (unsigned long long)a.ovf_cap;
        return __atomic_load_n(&a.queue[1], __ATOMIC_RELAXED) >= total;
    }
    if (a.sub_mode == 0)
        return (long long)__atomic_load_n(&a.queue[0], __ATOMIC_RELAXED) >= a.e_hi - a.e_lo;
    const unsigned long long made = __atomic_load_n(a.in_n, __ATOMIC_RELAXED);
    const unsigned long long total = made < (unsigned long long)a.in_cap ? made : (unsigned long long)a.in_cap;
    return __atomic_load_n(&a.queue[1 + a.sub_mode], __ATOMIC_RELAXED) >= total;
}
template <bool GS>
__device__ __forceinline__ bool want_split(const KArgs &a, const WaveState *S) {
    if (a.split_ticks <= 0 || a.sub_mode >= kSplitLevels) return false;
    if (uni64(wall_clock64()) - S->t0 <= a.split_ticks) return false;
    return queue_drained<GS>(a);
}

// Hands candidates [lo, hi) of the depth-d frame's node to the next split
// launch, in at most 64 chunks of at least min_chunk. False, with nothing
// handed on, when the wave's record of donations or the list is full
// (reserved entries past the end of the list are left empty). Donations are
// recorded for cancellation only where the donor can still abort (LDS
// buffers); the HBM launch's tasks cannot, so their donations are unlimited.
__device__ __noinline__ bool donate(const Ctx C, int d, int node, int lo, int hi, int min_chunk, bool record) {
    const KArgs &a = C.a;
    WaveState *S = C.S;
    if (record && S->n_don >= kMaxDonations) return false;
    const long long edge = S->sub_edge >= 0 ? S->sub_edge : S->rb + (S->f_cur[1] - 1);
    const int r = hi - lo;
    const int chunk = max(min_chunk, (r + W - 1) / W);
    const int parts = (r + chunk - 1) / chunk;
    unsigned long long b = 0;
    if (C.ln == 0) b = atomicAdd(a.out_n, (unsigned long long)parts);
    const long long base = uni64((long long)__shfl(b, 0));
    const long long room = a.out_cap - base;
    const int fit = room <= 0 ? 0 : (room < parts ? (int)room : parts);
    for (int k = C.ln; k < fit; k += W) {
        SplitTask *t = a.subs + a.out_base + base + k;
        t->edge = edge;
        t->node = node;
        t->depth = d;
        t->lo = fit == parts ? lo + k * chunk : 0;
        t->hi = fit == parts ? min(hi, lo + (k + 1) * chunk) : 0;
        for (int j = 2; j < d; j++) t->pos[j] = j < S->base_depth ? S->path_pos[j] : S->f_cur[j] - 1;
    }
    if (fit < parts) return false;
    if (record && C.ln == 0) {
        S->don_base[S->n_don] = a.out_base + base;
        S->don_cnt[S->n_don] = parts;
        S->n_don++;
    }
    wsync();
    return true;
}

// Donates the rest of the shallowest frame (depth >= base_depth, up to d)
// that has two or more candidates left; true if it did. The frame is closed
// by lowering its end to its cursor: f_cur[k] - 1 stays the position of the
// candidate being walked at depth k, which later donations from deeper
// frames record as their path.
template <bool GS>
__device__ __noinline__ bool split_shallowest(const Ctx C, int d) {
    WaveState *S = C.S;
    for (int k = uni(S->base_depth); k <= d; k++) {
        const int cur = uni(S->f_cur[k]), fh = uni(S->f_hi[k]);
        if (fh - cur < 2) continue;
        if (!donate(C, k, uni(S->f_node[k]), cur, fh, C.P.nodes[uni(S->f_node[k])].leaf ? W : 1, !GS)) return false;
        if (C.ln == 0) S->f_hi[k] = cur;
        wsync();
        return true;
    }
    return false;
}

// Value of one IEP factor. Per-lane candidate c sits at depth d; leaf
// counters are read for this lane.
template <bool GS>
__device__ __forceinline__ long long vid_value(const Ctx &C, const GatVid &v, int d, int c) {
    const WaveState *S = C.S;
    const GatSlot &s = C.P.slots[v.slot];
    if (s.leaf_idx >= 0) return S->leafcnt[s.leaf_idx][C.ln];
    const int L = S->slen[v.slot];
    if (v.mode == GATP_VID_COUNT) return L;
    return with_slot<GS>(C, v.slot, [&](auto p) -> long long {
        if (v.mode == GATP_VID_LIST_REF) {
            long long r = L;
            for (int i = 0; i < v.refn; i++) r -= has(p, L, i == d ? c : S->val[i], C.a.lin_thr) ? 1 : 0;
            return r;
        }
        const int t = v.bound ? mask_bound(S, v.bound, d, c) : INT_MAX;
        long long r = t == INT_MAX ? L : lbound(p, L, t);
        for (int j = 0; j <= d; j++)
            if (v.excl >> j & 1u) {
                const int x = j == d ? c : S->val[j];
                if (x < t && has(p, L, x, C.a.lin_thr)) r--;
            }
        return r;
    });
}

template <bool GS>
__device__ __forceinline__ unsigned long long term_value(const Ctx &C, const GatTerm &T, int d, int c) {
    unsigned long long local = 0, val = 1;
    int last = -1;
    for (int g = 0; g < T.sub_count; g++) {
        const GatSub &sb = C.P.subs[T.sub_begin + g];
        const unsigned long long x = (unsigned long long)vid_value<GS>(C, C.P.vids[sb.vid], d, c);
        val = (g == last + 1) ? x : val * x;
        if (sb.flag) {
            last = g;
            local += val * (unsigned long long)(long long)sb.coef;
        }
    }
    return local;
}

// replay: rebuild the node's sets for a split subtask without adding its
// terms again (the donor counted them).
template <bool GS>
__device__ __forceinline__ bool visit(const Ctx &C, const GatNode &N, int d, int c, bool replay = false) {
    WaveState *S = C.S;
    if (C.ln == 0) S->val[d] = c;
    wsync();
    const bool dup = N.ap_levels && equals_any(S, N.ap_levels, d, c);
    for (int k = 0; k < N.op_count; k++) {
        exec_op_uniform<GS>(C, C.P.ops[N.op_begin + k], d, c, dup);
        wsync();
        if (S->abort) return false;
    }
    for (int k = 0; k < N.term_count && !replay; k++) {
        const GatTerm &T = C.P.terms[N.term_begin + k];
        const unsigned long long v = term_value<GS>(C, T, d, c);
        if (C.ln == 0) S->task_acc[T.pattern] += v;
    }
    wsync();
    return true;
}

// All candidates of a leaf frame, flattened across lanes.
// With can_split, a wave that is due to split (want_split) hands the
// batches after the current one to the next split launch.
template <bool GS>
__device__ __forceinline__ void leaf(const Ctx &C, const GatNode &N, int d, int cur, int hi, bool can_split = false,
                                     int node = -1) {
    const KArgs &a = C.a;
    WaveState *S = C.S;
    const int ln = C.ln;
    unsigned long long bytes = 0;
    for (int cb = cur; cb < hi; cb += W) {
        const int ci = cb + ln;
        int c = 0;
        if (ci < hi) c = with_slot<GS>(C, N.loop_slot, [&](auto p) { return p[ci]; });
        // A chunk whose candidates' neighbour lists add up to more elements
        // than one wave sweeps within the split budget (about 3 ticks of the
        // 100 MHz clock per element) is not swept whole once the queue is
        // drained. First the rest of the shallowest loop above the leaf that
        // has candidates left goes to the next split launch (the most work
        // for one donation record). With none left, the leaf's candidates
        // after the first go, in about as many parts as the chunk holds
        // budgets of work (at most 64), so that whole waves sweep their lists
        // in parallel; the wave keeps the first candidate.
        if (can_split && hi - cb > 1 && queue_drained<GS>(a)) {
            const long long lw = ci < hi ? a.off[c + 1] - a.off[c] : 0;
            const long long tot = uni64((long long)wave_sum64((unsigned long long)lw));
            if (3 * tot > a.split_ticks && !split_shallowest<GS>(C, d - 1)) {
                const long long budgets = 3 * tot / a.split_ticks + 1;
                const long long np = budgets < W ? budgets : W;
                const int r = hi - cb - 1;
                if (donate(C, d, node, cb + 1, hi, (int)((r + np - 1) / np), !GS)) hi = cb + 1;
            }
        }
        const bool valid = ci < hi;
        bool pass = valid && !equals_any(S, N.excl, d, c);
        const bool dup = pass && N.ap_levels && equals_any(S, N.ap_levels, d, c);
        const long long ab = valid ? a.off[c] : 0;
        const int deg = valid ? (int)(a.off[c + 1] - ab) : 0;
        for (int k = 0; k < N.op_count; k++) {
            const GatOp &o = C.P.ops[N.op_begin + k];
            const int t = group_bound(S, o.groups, o.ngroups, d, c);
            int alen = (pass && !dup) ? trunc_len(a, c, ab, deg, t) : 0;
            if (o.kind == GATP_ALIAS) {
                S->leafcnt[o.leaf_idx][ln] = alen;
                continue;
            }
            // COUNT: |N(c) ∩ father ∩ [0,t)| minus excluded chosen vertices
            int ex[8];
            int nex = 0;
            for (int f = 0; f < d; f++)
                if (o.excl >> f & 1u) ex[nex++] = S->val[f];
            if (o.father < 0) {
                // |N(c) ∩ [0,t)| minus the excluded chosen vertices it holds
                const bool live = pass && !dup;
                S->leafcnt[o.leaf_idx][ln] = live ? alen - excluded_in_list(a, c, ab, alen, ex, nex) : 0;
                if (a.account && live) bytes += 4ull * alen + 16ull;
                continue;
            }
            const int f = o.father;
            const bool has_r = f >= 0;
            bool b_lds = false;
            int blen = 0;
            // Bg: father in HBM (a neighbour list, or any set in the overflow
            // launch); Bl: father in this wave's LDS buffers. hub_b: the
            // father is a hub's neighbour list in HBM, probed by bitmap.
            const int *Bg = a.adj;
            const int *Bl = C.store;
            int hub_b = -1, lim_b = 0;
            if (has_r) {
                const GatSlot &fs = C.P.slots[f];
                b_lds = in_lds<GS>(C, f);
                if (b_lds) Bl = C.store + (long long)fs.store * a.cap;
                else if (fs.kind == GATP_ALIAS) Bg = a.adj + S->sbeg[f];
                else if constexpr (GS) Bg = C.store + (long long)fs.store * a.cap;
                const int flen = S->slen[f];
                if (pass && !dup) blen = t == INT_MAX ? flen : (b_lds ? lbound(Bl, flen, t) : lbound(Bg, flen, t));
                if (!b_lds && fs.kind == GATP_ALIAS) hub_b = uni(hub_of(a, S->val[fs.depth]));
            }
            // candidate lists with a hub bitmap are probed through it
            const int hub_a = (has_r && valid) ? hub_of(a, c) : -1;
            const int lim_a = hub_a >= 0 ? list_lim(a, ab, alen) : 0;
            if (hub_b >= 0 && pass && !dup) lim_b = blen > 0 ? Bg[blen - 1] : -1;
            const bool sa = !has_r || stream_a(a, alen, blen, b_lds || hub_b >= 0, hub_a >= 0);
            const int mylen = sa ? alen : blen;
            int T = 0;
            const int myoff = excl_scan(mylen, &T);
            int cnt = 0;
            for (int k0 = 0; k0 < T; k0 += W) {
                const int j = k0 + ln;
                bool hit = false;
                by_owner(myoff, k0, T, [&](const auto &pull) {
                    const int s_off = pull(myoff);
                    const int s_sa = pull((int)sa);
                    const long long s_ab = pull(ab);
                    const int s_alen = pull(alen);
                    const int s_blen = pull(blen);
                    const int s_hub = pull(hub_a);
                    const int s_lim = pull(lim_a);
                    const int s_limb = pull(lim_b);
                    if (j < T) {
                        const int idx = j - s_off;
                        int x;
                        if (s_sa) {
                            x = a.adj[s_ab + idx];
                            if (!has_r) hit = true;
                            else if (b_lds) hit = has(Bl, s_blen, x, a.lin_thr);
                            else if (hub_b >= 0) hit = hub_has(a, hub_b, x, s_limb);
                            else hit = has(Bg, s_blen, x, a.lin_thr);
                        } else {
                            x = b_lds ? Bl[idx] : Bg[idx];
                            hit = s_hub >= 0 ? hub_has(a, s_hub, x, s_lim) : has(a.adj + s_ab, s_alen, x, a.lin_thr);
                        }
                        for (int q = 0; q < nex; q++) hit &= (x != ex[q]);
                    }
                });
                const unsigned long long m = ballot(hit);
                const int lo = max(myoff - k0, 0), hi2 = min(myoff + mylen - k0, W);
                if (lo < hi2) cnt += __popcll(m & range_mask(lo, hi2));
            }
            S->leafcnt[o.leaf_idx][ln] = cnt;
            if (a.account && pass && !dup)
                bytes += 4ull * (alen + blen) + 16ull +
                         ((has_r && C.P.slots[f].kind == GATP_ALIAS) ? 16ull : 0ull);
        }
        wsync();
        for (int k = 0; k < N.term_count; k++) {
            const GatTerm &Tm = C.P.terms[N.term_begin + k];
            unsigned long long v = pass ? term_value<GS>(C, Tm, d, c) : 0ull;
            v = wave_sum64(v);
            if (ln == 0) S->task_acc[Tm.pattern] += v;
        }
        wsync();
        if (can_split && hi - (cb + W) >= 2 * W && want_split<GS>(a, S) && donate(C, d, node, cb + W, hi, W, !GS))
            break;
    }
    if (a.account) {
        bytes = wave_sum64(bytes);
        if (ln == 0) S->ebytes += bytes;
    }
}

template <bool GS>
__device__ __forceinline__ void push_frame(const Ctx &C, int d, int node, int lo, int hi) {
    WaveState *S = C.S;
    const GatNode &N = C.P.nodes[node];
    const int ls = N.loop_slot;
    const int L = S->slen[ls];
    int hb = L;
    if (N.bound) {
        const int t = mask_bound(S, N.bound, d, 0);  // bound levels are all < d
        hb = uni(with_slot<GS>(C, ls, [&](auto p) { return lbound(p, L, t); }));
    }
    if (C.ln == 0) {
        S->f_node[d] = node;
        S->f_cur[d] = max(lo, 0);
        S->f_hi[d] = min(hi, hb);
    }
    wsync();
}

// Pending counts and W_alg of the edges finished since the last commit go
// into the wave's totals; an abandoned edge's are dropped instead.
__device__ __forceinline__ void commit_pending(WaveState *S, int np, int ln) {
    for (int i = ln; i < np; i += W) {
        S->wave_acc[i] += S->task_acc[i];
        S->task_acc[i] = 0;
    }
    if (ln == 0) {
        S->bytes += S->ebytes;
        S->ebytes = 0;
        S->n_don = 0;  // what this work handed on is final too
    }
    wsync();
}

__device__ __forceinline__ void drop_pending(WaveState *S, int np, int ln) {
    for (int i = ln; i < np; i += W) S->task_acc[i] = 0;
    if (ln == 0) {
        S->ebytes = 0;
        S->abort = 0;
    }
    wsync();
}

// Depth-first walk of the trie for the directed edges [e0, e1). When the
// root has one child (a single depth-1 node, so each edge is one subtree),
// an edge whose sets outgrow the LDS buffers is handed alone to the HBM
// launch and the walk goes on with the next edge; otherwise the whole task
// is abandoned and re-run there.
template <bool GS>
__device__ __forceinline__ void run_task(const Ctx &C, long long e0, long long e1) {
    const KArgs &a = C.a;
    const GatPlanDev &P = C.P;
    WaveState *S = C.S;
    const GatNode &root = P.nodes[0];
    const bool edge_grain = !GS && root.child_count == 1;
    long long e = e0;
    while (e < e1) {
        // row of edge e: last v with off[v] <= e
        int lo = 0, hi = a.n;  // invariant off[lo] <= e < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.off[mid] <= e) lo = mid; else hi = mid;
        }
        const int v0 = uni(lo);
        const long long rb = a.off[v0], re = a.off[v0 + 1];
        const long long se = e1 < re ? e1 : re;
        if (C.ln == 0) {
            S->val[0] = v0;
            S->rb = rb;
            S->sub_edge = -1;
            S->base_depth = 2;
        }
        wsync();
        for (int k = 0; k < root.op_count; k++) exec_op_uniform<GS>(C, P.ops[root.op_begin + k], 0, v0, false);
        wsync();
        const int lo1 = (int)(e - rb), hi1 = (int)(se - rb);
        int d = 1;
        push_frame<GS>(C, 1, root.child_begin, lo1, hi1);
        while (d >= 1) {
            const int node = uni(S->f_node[d]);
            const GatNode &N = P.nodes[node];
            bool exhausted = false;
            // loops at depth >= 2 may be handed to a split launch
            const bool can_split = d >= 2;
            if (N.leaf) {
                leaf<GS>(C, N, d, uni(S->f_cur[d]), uni(S->f_hi[d]), can_split, node);
                exhausted = true;
            } else {
                const int cur = uni(S->f_cur[d]), fh = uni(S->f_hi[d]);
                if (can_split && cur < fh && want_split<GS>(a, S) && split_shallowest<GS>(C, d)) continue;
                if (cur < fh) {
                    const int c = uni(with_slot<GS>(C, N.loop_slot, [&](auto p) { return p[cur]; }));
                    if (C.ln == 0) S->f_cur[d] = cur + 1;
                    wsync();
                    if (N.excl && equals_any(S, N.excl, d, c)) continue;
                    if (edge_grain && d == 1) commit_pending(S, P.n_patterns, C.ln);  // previous edge done
                    if (!visit<GS>(C, N, d, c)) {  // a set outgrew its LDS buffer
                        if (!edge_grain) return;  // the whole task goes to the HBM launch
                        if (C.ln == 0) {
                            const unsigned long long idx = atomicAdd(a.ovf_n, 1ull);
                            const long long edge = rb + (uni(S->f_cur[1]) - 1);
                            if ((long long)idx < a.ovf_cap) a.ovf[idx] = edge; else atomicOr(a.err, 2u);
                        }
                        cancel_donations(a, S, C.ln);
                        drop_pending(S, P.n_patterns, C.ln);
                        d = 1;
                        continue;
                    }
                    if (N.child_count > 0) {
                        d++;
                        push_frame<GS>(C, d, N.child_begin, 0, INT_MAX);
                    }
                    continue;
                }
                exhausted = true;
            }
            if (exhausted) {
                if (N.next_sibling >= 0)
                    push_frame<GS>(C, d, N.next_sibling, d == 1 ? lo1 : 0, d == 1 ? hi1 : INT_MAX);
                else
                    d--;
            }
        }
        e = se;
    }
    if (edge_grain) commit_pending(S, P.n_patterns, C.ln);
}

// Split subtask *g (in subs[]): the sets of the root and of the nodes on its
// path are rebuilt (neither counted nor accounted again), then the walk
// covers candidates [lo, hi) of its node and their subtrees only. A set that
// outgrows its LDS buffer leaves S->abort set (the HBM launch re-runs it).
template <bool GS>
__device__ __forceinline__ void run_sub(const Ctx &C, const SplitTask *g) {
    const KArgs &a = C.a;
    const GatPlanDev &P = C.P;
    WaveState *S = C.S;
    const long long e = uni64(g->edge);
    const int sub_node = uni(g->node), D = uni(g->depth), sub_lo = uni(g->lo), sub_hi = uni(g->hi);
    int lo = 0, hi = a.n;  // off[lo] <= e < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.off[mid] <= e) lo = mid; else hi = mid;
    }
    const int v0 = uni(lo);
    const long long rb = a.off[v0];
    const unsigned long long eb = S->ebytes;
    if (C.ln == 0) {
        S->val[0] = v0;
        S->rb = rb;
        S->sub_edge = e;
        S->base_depth = D;
    }
    for (int j = 2 + C.ln; j < D; j += W) S->path_pos[j] = g->pos[j];
    wsync();
    const GatNode &root = P.nodes[0];
    for (int k = 0; k < root.op_count; k++) exec_op_uniform<GS>(C, P.ops[root.op_begin + k], 0, v0, false);
    wsync();
    bool ok = true;
    for (int j = 1; j < D && ok; j++) {
        int nj = sub_node;  // the ancestor of the subtask's node at depth j
        for (int t = D; t > j; t--) nj = P.nodes[nj].parent;
        const GatNode &Nj = P.nodes[nj];
        const int pos = j == 1 ? (int)(e - rb) : uni(S->path_pos[j]);
        const int vj = uni(with_slot<GS>(C, Nj.loop_slot, [&](auto p) { return p[pos]; }));
        ok = visit<GS>(C, Nj, j, vj, true);
    }
    if (C.ln == 0) S->ebytes = eb;
    wsync();
    if (!ok) return;
    const bool can_split = !GS && a.sub_mode < kSplitLevels;
    push_frame<GS>(C, D, sub_node, sub_lo, sub_hi);
    int d = D;
    while (d >= D) {
        const int node = uni(S->f_node[d]);
        const GatNode &N = P.nodes[node];
        bool exhausted = false;
        if (N.leaf) {
            leaf<GS>(C, N, d, uni(S->f_cur[d]), uni(S->f_hi[d]), can_split, node);
            exhausted = true;
        } else {
            const int cur = uni(S->f_cur[d]), fh = uni(S->f_hi[d]);
            if (can_split && cur < fh && want_split<GS>(a, S) && split_shallowest<GS>(C, d)) continue;
            if (cur < fh) {
                const int c = uni(with_slot<GS>(C, N.loop_slot, [&](auto p) { return p[cur]; }));
                if (C.ln == 0) S->f_cur[d] = cur + 1;
                wsync();
                if (N.excl && equals_any(S, N.excl, d, c)) continue;
                if (!visit<GS>(C, N, d, c)) return;
                if (N.child_count > 0) {
                    d++;
                    push_frame<GS>(C, d, N.child_begin, 0, INT_MAX);
                }
                continue;
            }
            exhausted = true;
        }
        if (exhausted) {
            if (d > D && N.next_sibling >= 0)
                push_frame<GS>(C, d, N.next_sibling, 0, INT_MAX);
            else
                d--;
        }
    }
}

// The array-store kernels read their arguments from device memory (written
// by put_kargs just before the launch): the walk hands a reference to them
// to out-of-line helpers, and a reference to a by-value kernel argument
// would make every lane copy the whole record to scratch first.
__global__ void put_kargs(KArgs a, KArgs *dst) {
    if (threadIdx.x == 0) *dst = a;
}

template <bool GS>
__global__ void __launch_bounds__(256) engine_kernel(const KArgs *__restrict__ ap, int wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const KArgs &a = *ap;
    const int wid = threadIdx.x / W, ln = threadIdx.x % W;
    const int wpb = blockDim.x / W;
    WaveState *S = reinterpret_cast<WaveState *>(smem + (size_t)wid * wave_bytes);
    const GatPlanDev &P = *a.plan;
    int *store;
    if (GS)
        store = a.gstore + (long long)(blockIdx.x * wpb + wid) * P.n_store * (long long)a.cap;
    else
        store = reinterpret_cast<int *>(smem + (size_t)wid * wave_bytes + sizeof(WaveState));
    Ctx C{a, P, S, store, ln};
    const int np = P.n_patterns;
    const long long tl_t0 = wall_clock64();
    long long tl_items = 0;
    for (int i = ln; i < np; i += W) S->wave_acc[i] = 0;
    if (ln == 0) {
        S->bytes = 0;
        S->ebytes = 0;
    }
    wsync();
    for (;;) {
        long long e0;  // the task: directed edges [e0, e0 + chunk)
        if (GS) {
            unsigned long long r = 0;
            if (ln == 0) r = atomicAdd(&a.queue[1], 1ull);
            r = (unsigned long long)uni64((long long)__shfl(r, 0));
            const unsigned long long total = *a.ovf_n;
            if (r >= total || (long long)r >= a.ovf_cap) break;
            e0 = uni64(a.ovf[r]);
        } else {
            unsigned long long r = 0;
            if (ln == 0) r = atomicAdd(&a.queue[0], (unsigned long long)a.chunk);
            e0 = a.e_lo + uni64((long long)__shfl(r, 0));
            if (e0 >= a.e_hi) break;
        }
        tl_items++;
        for (int i = ln; i < np; i += W) S->task_acc[i] = 0;
        if (ln == 0) {
            S->abort = 0;
            S->n_don = 0;
            S->t0 = wall_clock64();
        }
        wsync();
        run_task<GS>(C, e0, e0 + a.chunk < a.e_hi ? e0 + a.chunk : a.e_hi);
        wsync();
        if (S->abort) {
            if (GS) {
                if (ln == 0) atomicOr(a.err, 1u);  // cannot happen: HBM buffers hold max degree
            } else {
                cancel_donations(a, S, ln);
                if (ln == 0) {
                    const unsigned long long idx = atomicAdd(a.ovf_n, 1ull);
                    if ((long long)idx < a.ovf_cap) a.ovf[idx] = e0; else atomicOr(a.err, 2u);
                }
            }
            if (ln == 0) S->ebytes = 0;  // the HBM launch does this task again
        } else {
            for (int i = ln; i < np; i += W) S->wave_acc[i] += S->task_acc[i];
            if (ln == 0) {
                S->bytes += S->ebytes;
                S->ebytes = 0;
            }
        }
        wsync();
    }
    for (int i = ln; i < np; i += W)
        if (S->wave_acc[i]) atomicAdd(&a.counts[i], S->wave_acc[i]);
    if (a.account && ln == 0 && S->bytes) atomicAdd(a.bytes, S->bytes);
    if (a.tl && ln == 0)
        tl_write(a.tl, a.tl_base + blockIdx.x * wpb + wid, GS ? kLevelsHbm : 0, blockIdx.x * wpb + wid, tl_t0,
                 wall_clock64(), 0, 0, tl_items, 0);
}

// Subtask launches: list sub_mode (1 .. kSplitLevels: LDS buffers; then the
// HBM list with HBM buffers). A subtask that outgrows LDS is copied to the
// HBM list for the last launch.
template <bool GS>
__global__ void __launch_bounds__(256) engine_sub_kernel(const KArgs *__restrict__ ap, int wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const KArgs &a = *ap;
    const int wid = threadIdx.x / W, ln = threadIdx.x % W;
    const int wpb = blockDim.x / W;
    WaveState *S = reinterpret_cast<WaveState *>(smem + (size_t)wid * wave_bytes);
    const GatPlanDev &P = *a.plan;
    int *store;
    if (GS)
        store = a.gstore + (long long)(blockIdx.x * wpb + wid) * P.n_store * (long long)a.cap;
    else
        store = reinterpret_cast<int *>(smem + (size_t)wid * wave_bytes + sizeof(WaveState));
    Ctx C{a, P, S, store, ln};
    const int np = P.n_patterns;
    const long long tl_t0 = wall_clock64();
    long long tl_items = 0;
    for (int i = ln; i < np; i += W) S->wave_acc[i] = 0;
    if (ln == 0) {
        S->bytes = 0;
        S->ebytes = 0;
    }
    wsync();
    for (;;) {
        unsigned long long r = 0;
        if (ln == 0) r = atomicAdd(&a.queue[1 + a.sub_mode], 1ull);
        r = (unsigned long long)uni64((long long)__shfl(r, 0));
        const unsigned long long made = *a.in_n;
        const unsigned long long total = made < (unsigned long long)a.in_cap ? made : (unsigned long long)a.in_cap;
        if (r >= total) break;
        const SplitTask *g = a.subs + a.in_base + (long long)r;
        tl_items++;
        for (int i = ln; i < np; i += W) S->task_acc[i] = 0;
        if (ln == 0) {
            S->abort = 0;
            S->n_don = 0;
            S->t0 = wall_clock64();
        }
        wsync();
        if (uni(g->lo) < uni(g->hi)) run_sub<GS>(C, g);
        wsync();
        if (S->abort) {
            if (GS) {
                if (ln == 0) atomicOr(a.err, 1u);  // cannot happen: HBM buffers hold max degree
            } else {
                cancel_donations(a, S, ln);
                unsigned long long idx = 0;
                if (ln == 0) idx = atomicAdd(a.sovf_n, 1ull);
                idx = (unsigned long long)uni64((long long)__shfl(idx, 0));
                if ((long long)idx < a.sovf_cap) {
                    SplitTask *t = a.subs + a.sovf_base + idx;
                    for (int k = ln; k < (int)(sizeof(SplitTask) / sizeof(int)); k += W)
                        reinterpret_cast<int *>(t)[k] = reinterpret_cast<const int *>(g)[k];
                } else if (ln == 0) {
                    atomicOr(a.err, 2u);
                }
                if (ln == 0) atomicAdd(&a.ovf_n[2], 1ull);  // statistics
            }
            if (ln == 0) S->ebytes = 0;
        } else {
            for (int i = ln; i < np; i += W) S->wave_acc[i] += S->task_acc[i];
            if (ln == 0) {
                S->bytes += S->ebytes;
                S->ebytes = 0;
            }
        }
        wsync();
    }
    for (int i = ln; i < np; i += W)
        if (S->wave_acc[i]) atomicAdd(&a.counts[i], S->wave_acc[i]);
    if (a.account && ln == 0 && S->bytes) atomicAdd(a.bytes, S->bytes);
    if (a.tl && ln == 0)
        tl_write(a.tl, a.tl_base + blockIdx.x * wpb + wid, a.sub_mode + (GS ? 1 : 0), blockIdx.x * wpb + wid, tl_t0,
                 wall_clock64(), 0, 0, tl_items, 0);
}

// ---------------------------------------------------------------------------
// Bitmap vertex store for clique plans (the reference declares
// VertexStoreType::Bitmap in include/configs/vertex_set_config.hpp:3-6 but
// never implements it). On the degree-oriented graph every k-clique is a
// chain v0 > v1 > ... with v_{i+1} in N+(v_i). Per vertex v0 the wave stages
// L = N+(v0) (sorted, d entries) and builds the oriented adjacency of the
// subgraph induced on L as bit rows over positions in L:
//   M_p = { q < p : L[q] in N+(L[p]) }.
// Every clique with top vertex v0 is then a clique of these rows:
//   #K3 += sum_p popc(M_p)
//   #K4 += sum_p sum_{q in M_p} popc(M_p & M_q)
//   #K5 += sum over q in M_p, r in M_p & M_q of popc(M_p & M_q & M_r)
//   #K6 one level deeper.
// Each neighbour pair inside N+(v0) is tested once per v0 (one flattened
// sweep over the lists N+(L[p]), binary search in L or the reverse for
// skewed pairs), so the deeper levels are pure bit operations. Tiers by d:
// the main launch gives each vertex with d <= 256 one wave and 8-word rows in
// LDS; larger vertices are shared by the 8 waves of a workgroup, with
// 16-word rows in LDS (d <= 512) or rows in the workgroup's HBM scratch.
constexpr int CMAXW = 8;   // main tier: d <= 256
constexpr int CMIDW = 16;  // workgroup tier 0: d <= 512, rows in LDS
constexpr int CBIGW = 32;  // workgroup tier 1: d <= 1024, rows in LDS when it fits, else HBM
// Workgroup-tier items are (vertex, part) pairs, vertex * CSPLIT_ENC + part:
// the parts of a vertex share out its pairs (each builds the rows).
constexpr int CSPLIT_ENC = 16;
constexpr int CSPLIT12 = 2;   // parts per vertex in tiers 1 and 2
constexpr int CSPLIT0 = 1;    // parts per tier-0 vertex with d > CSPLIT0_D
constexpr int CSPLIT0_D = 256;
__device__ __forceinline__ int clique_parts(int tier, int d) {
    return tier ? CSPLIT12 : (d > CSPLIT0_D ? CSPLIT0 : 1);
}
constexpr int CBLOCKW0 = 8;   // waves per workgroup, tier 0
constexpr int CBLOCKW1 = 16;  // waves per workgroup, tiers 1 and 2 (few, heavy vertices)
// Words per bit row of a list of d vertices: enough for d bits, rounded up
// to an odd number so that lanes reading the same word of different rows
// hit different LDS banks (an even stride folds them onto a few banks).
__host__ __device__ constexpr int row_words(int d) { return ((d + 31) >> 5) | 1; }
// LDS index of a workgroup tier's list L (lists of up to lcap vertices):
// open addressing with linear probing over 2 lcap 16-bit slots (a power of
// two), each 0 or 1 + the position in L of the vertex stored there.
__host__ __device__ constexpr int index_slots(int lcap) { return 2 * lcap; }
__device__ __forceinline__ unsigned index_hash(int y, unsigned mask) {
    return ((unsigned)y * 0x9E3779B1u >> 16) & mask;
}

struct CArgs {
    const long long *off;
    const int *adj;
    const int *split;
    int n;
    long long m;
    int root_lo, root_hi;        // top vertices of this count: [root_lo, root_hi)
    unsigned long long *counts;
    int pat_of_k[8];
    int kmax;
    unsigned long long *queue;   // [0] main vertex queue, [1 + t] workgroup tier t
    long long *tier[3];          // vertices of the workgroup tiers (counts ovf_n[t])
    unsigned long long *ovf_n;
    long long ovf_cap;
    unsigned long long *bytes;   // W_alg accumulator of this launch
    unsigned int *err;
    unsigned *grows;             // HBM tier: per-workgroup rows, hcap * ceil(hcap/32) words
    int *goff;                   // HBM tier: per-workgroup pair offsets, hcap + 1 entries
    int hcap;                    // HBM tier capacity (>= maximum out-degree)
    int account;
    int cap;                     // main-tier limit on d (<= 32 * CMAXW)
    int chunk;                   // vertices per main-queue grab
    int lin_thr;
    int gallop_ratio;
    long long *tl;               // timeline records (GAT_FLAG_TIMELINE), or null
    int tl_base;                 // first record of this launch
};

// Rows M_p of the list L (d entries) into rows (w words per row, already
// zeroed), for the rows p = first, first + stride, first + 2 stride, ...
// (a workgroup's waves interleave their rows, so light and heavy rows mix):
// lanes sweep the (row, element) pairs of 64 of these rows at a time; each
// element of N+(L[p]) is searched in L[0,p) (by the LDS index ix of L when
// given, else by binary search), or each L[q], q < p, in N+(L[p]) when that
// list is gallop_ratio times longer than p.
template <bool ACC, class LP, class RP>
__device__ void vtx_rows(const CArgs &a, LP L, int d, RP rows, int w, int first, int stride, int ln,
                         unsigned long long &bytes, const unsigned short *ix = nullptr, unsigned ix_mask = 0) {
    for (int cb = 0; first + cb * stride < d; cb += W) {
        const int ci = first + (cb + ln) * stride;
        const bool valid = ci < d;
        const int x = valid ? L[ci] : 0;
        const long long xb = valid ? a.off[x] : 0;
        const int xl = valid ? a.split[x] : 0;  // N+(x): all below x
        const bool sx = (long long)xl <= (long long)a.gallop_ratio * (long long)ci;
        const int mylen = valid ? (sx ? xl : ci) : 0;
        if (ACC && valid) bytes += 4ull * (xl + ci) + 16ull;
        int T = 0;
        const int myoff = excl_scan(mylen, &T);
        for (int k0 = 0; k0 < T; k0 += W) {
            const int j = k0 + ln;
            // every lane takes part in the cross-lane reads: a segment's
            // owner lane may itself have no element here
            by_owner(myoff, k0, T, [&](const auto &pull) {
                const int s_off = pull(myoff);
                const int s_sx = pull((int)sx);
                const long long s_xb = pull(xb);
                const int s_xl = pull(xl);
                const int s_ci = pull(ci);
                if (j < T) {
                    const int idx = j - s_off;
                    int pos = -1;
                    if (s_sx) {
                        const int y = a.adj[s_xb + idx];
                        if (ix) {  // probe the index of L: a slot of 0 ends the probe
                            for (unsigned h = index_hash(y, ix_mask);; h = (h + 1) & ix_mask) {
                                const int e = ix[h];
                                if (e == 0) break;
                                if (L[e - 1] == y) {
                                    if (e - 1 < s_ci) pos = e - 1;
                                    break;
                                }
                            }
                        } else {
                            const int p = lbound(L, s_ci, y);
                            if (p < s_ci && L[p] == y) pos = p;
                        }
                    } else {
                        if (has(a.adj + s_xb, s_xl, L[idx], a.lin_thr)) pos = idx;
                    }
                    if (pos >= 0) atomicOr(&rows[(long long)s_ci * w + (pos >> 5)], 1u << (pos & 31));
                }
            });
        }
    }
}

// Position of the j-th set bit (0-based) of x, j < popc(x).
__device__ __forceinline__ int select_bit(unsigned x, int j) {
    int pos = 0, c;
    c = __popc(x & 0xffffu);
    if (j >= c) { j -= c; x >>= 16; pos += 16; }
    c = __popc(x & 0xffu);
    if (j >= c) { j -= c; x >>= 8; pos += 8; }
    c = __popc(x & 0xfu);
    if (j >= c) { j -= c; x >>= 4; pos += 4; }
    c = __popc(x & 0x3u);
    if (j >= c) { j -= c; x >>= 2; pos += 2; }
    if (j >= (int)(x & 1u)) pos += 1;
    return pos;
}

// popc(M_p): row p has bits below p only.
template <class RP>
__device__ __forceinline__ int row_popc(RP rows, int p, int w) {
    int c = 0;
    for (int u = 0; u <= (p >> 5); u++) c += __popc(rows[(long long)p * w + u]);
    return c;
}

// t[i] for a run-time i < N, by selects (constant indexes keep t in registers).
template <int N>
__device__ __forceinline__ unsigned pick(const unsigned (&t)[N], int i) {
    unsigned v = t[0];
#pragma unroll
    for (int k = 1; k < N; k++) v = i == k ? t[k] : v;
    return v;
}

// sum over TC <= u < n of popc(a[u] & b[u]) (and of a[u] & b[u] & c[u]):
// with MAXW > 0 (n <= MAXW) the loop is unrolled, so every word is read at
// an immediate offset from its row's base and costs its reads, one bit op
// and one popcount; MAXW = 0 loops (rows of any width).
template <int MAXW, int TC, class RP>
__device__ __forceinline__ unsigned and2_popc(RP a, RP b, int n) {
    unsigned c = 0;
    if constexpr (MAXW > 0) {
#pragma unroll
        for (int u = TC; u < MAXW; u++) {
            if (u >= n) break;
            c += __popc(a[u] & b[u]);
        }
    } else {
        for (int u = TC; u < n; u++) c += __popc(a[u] & b[u]);
    }
    return c;
}
template <int MAXW, int TC, class RP>
__device__ __forceinline__ unsigned and3_popc(RP a, RP b, RP c3, int n) {
    unsigned c = 0;
    if constexpr (MAXW > 0) {
#pragma unroll
        for (int u = TC; u < MAXW; u++) {
            if (u >= n) break;
            c += __popc(a[u] & b[u] & c3[u]);
        }
    } else {
        for (int u = TC; u < n; u++) c += __popc(a[u] & b[u] & c3[u]);
    }
    return c;
}

// Counts at and below the pair level. The pairs (p, q), q in M_p, are
// numbered through off[p] = sum of popc(M_p') over p' < p (off[d] = P), and
// worker k takes pairs first, first + step, ..., so lanes share the work
// evenly however uneven the rows are. For a pair: t = M_p & M_q gives #K4
// += popc(t); each r in t adds popc(t & M_r) to #K5; each s in t & M_r adds
// popc(t & M_r & M_s) to #K6. Words are read from the rows as needed (no
// per-lane row copies: that keeps register use, and so occupancy, low),
// each loop stopping at the last word that can hold bits. Rows are w words
// apart (w <= MAXW when MAXW > 0); row offsets fit 32 bits and their
// factors 24 bits, so they take full-rate 24-bit multiplies.
template <int MAXW, bool ACC, class RP>
__device__ void vtx_pairs(const CArgs &a, RP rows, int d, int w, const int *off, int P, int first, int step,
                          unsigned long long *acc, unsigned long long &bytes) {
    for (int k = first; k < P; k += step) {
        int lo = 0, hi = d;  // off[lo] <= k < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= k) lo = mid; else hi = mid;
        }
        const int p = lo;
        const RP rp = rows + __umul24(p, w);
        int j = k - off[p];
        int q = 0;
        for (int u = 0; u <= (p >> 5); u++) {
            const unsigned x = rp[u];
            const int c = __popc(x);
            if (j < c) {
                q = u * 32 + select_bit(x, j);
                break;
            }
            j -= c;
        }
        const int wq = (q >> 5) + 1;
        if (ACC) bytes += 8ull * wq;
        const RP rq = rows + __umul24(q, w);
        // the first TC words of t = M_p & M_q stay in registers; the rest
        // are re-read from the rows when needed
        constexpr int TC = 2;
        unsigned tc[TC];
        unsigned c4 = 0;
#pragma unroll
        for (int u = 0; u < TC; u++) {
            tc[u] = u < wq ? (rp[u] & rq[u]) : 0u;
            c4 += __popc(tc[u]);
        }
        c4 += and2_popc<MAXW, TC>(rp, rq, wq);
        acc[4] += c4;
        if (a.kmax < 5) continue;
        unsigned c5 = 0;  // at most |t| * d bits: fits 32 bits
        if (a.kmax == 5) {
            // the next r's row words are read before the current r's are
            // used, so the LDS latency overlaps the bit operations
            for (int rw = 0; rw < wq; rw++) {
                unsigned br = rw < TC ? pick(tc, rw) : (rp[rw] & rq[rw]);
                if (!br) continue;
                int r = rw * 32 + __ffs(br) - 1;
                br &= br - 1;
                RP rr = rows + __umul24(r, w);
                unsigned x0 = rr[0], x1 = rr[1];
                for (;;) {
                    const bool more = br != 0;
                    const int rn = more ? rw * 32 + __ffs(br) - 1 : r;
                    br &= br - 1;
                    const RP rnp = rows + __umul24(rn, w);
                    const unsigned y0 = rnp[0], y1 = rnp[1];
                    if (ACC) bytes += 8ull * (rw + 1);
                    c5 += __popc(tc[0] & x0) + __popc(tc[1] & x1);
                    if (rw >= TC) c5 += and3_popc<MAXW, TC>(rp, rq, rr, rw + 1);
                    if (!more) break;
                    rr = rnp;
                    x0 = y0;
                    x1 = y1;
                }
            }
            acc[5] += c5;
            continue;
        }
        for (int rw = 0; rw < wq; rw++) {
            // every r in word rw of t needs words 0 .. rw (the same count
            // for all lanes still in the loop)
            unsigned br = rw < TC ? pick(tc, rw) : (rp[rw] & rq[rw]);
            while (br) {
                const int r = rw * 32 + __ffs(br) - 1;
                br &= br - 1;
                const RP rr = rows + __umul24(r, w);
                if (ACC) bytes += 8ull * (rw + 1);
                // words 0 and 1 of t sit in registers. Word 1 of M_r is read
                // whatever rw is: it is 0 when r < 32, and it lies within the
                // rows (lists of more than 32 vertices have rows of 3 or more
                // words; a shorter list has t's word 1 = 0, and its slice of
                // at least 32 words runs past row r + 1)
                c5 += __popc(tc[0] & rr[0]) + __popc(tc[1] & rr[1]);
                if (rw >= TC) c5 += and3_popc<MAXW, TC>(rp, rq, rr, rw + 1);
                if (a.kmax < 6) continue;
                unsigned c6 = 0;
                for (int sw = 0; sw <= rw; sw++) {
                    unsigned bs = rp[sw] & rq[sw] & rr[sw];
                    while (bs) {
                        const int sx = sw * 32 + __ffs(bs) - 1;
                        bs &= bs - 1;
                        const RP rs = rows + __umul24(sx, w);
                        if (ACC) bytes += 8ull * ((sx >> 5) + 1);
                        for (int u = 0; u <= (sx >> 5); u++) c6 += __popc(rp[u] & rq[u] & rr[u] & rs[u]);
                    }
                }
                acc[6] += c6;
            }
        }
        acc[5] += c5;
    }
}

template <bool ACC>
__device__ __forceinline__ void clique_flush(const CArgs &a, unsigned long long *acc, unsigned long long bytes,
                                             int ln) {
    for (int k = 3; k <= 6; k++) acc[k] = wave_sum64(acc[k]);
    if (ACC) bytes = wave_sum64(bytes);
    if (ln == 0) {
        for (int k = 3; k <= 6; k++)
            if (a.pat_of_k[k] >= 0 && acc[k]) atomicAdd(&a.counts[a.pat_of_k[k]], acc[k]);
        if (ACC && bytes) atomicAdd(a.bytes, bytes);
    }
}

// Main tier: one wave per vertex, vertices from the main queue, d <= a.cap
// (rows in LDS, 8 words wide); larger vertices go to the block tiers.
// MODE bit 0: accumulate W_alg; bit 1: timeline records (separate
// instantiations, so that a timed count carries none of their work or
// registers).
template <int MODE>
__global__ void __launch_bounds__(256) clique_kernel(CArgs a, int wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / W, ln = threadIdx.x % W;
    unsigned char *base = smem + (size_t)wid * wave_bytes;
    int *L = reinterpret_cast<int *>(base);
    unsigned *rows = reinterpret_cast<unsigned *>(L + a.cap + 4);  // L: d entries, then d + 1 offsets
    unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0}, bytes = 0;  // per lane
    constexpr bool ACC = MODE & 1, tl = (MODE & 2) != 0;
    const long long tl_t0 = tl ? wall_clock64() : 0;
    long long tl_build = 0, tl_count = 0, tl_items = 0, tl_pairs = 0;
    for (;;) {
        unsigned long long r = 0;
        if (ln == 0) r = atomicAdd(&a.queue[0], (unsigned long long)a.chunk);
        const long long v_lo = a.root_lo + uni64((long long)__shfl(r, 0));
        if (v_lo >= a.root_hi) break;
        const long long v_hi = v_lo + a.chunk < a.root_hi ? v_lo + a.chunk : a.root_hi;
        for (long long vv = v_lo; vv < v_hi; vv++) {
            const int v0 = (int)vv;
            const int d = uni(a.split[v0]);
            if (d < 2) continue;
            if (d > a.cap) {
                if (ln == 0) {
                    const int t = d <= 32 * CMIDW ? 0 : (d <= 32 * CBIGW ? 1 : 2);
                    const int parts = clique_parts(t, d);
                    atomicAdd(&a.ovf_n[3], 1ull);  // vertices handed on (statistics)
                    if (t == 2 && !a.grows) {
                        atomicOr(a.err, 16u);  // no HBM tier in this count (allocation-free entry point)
                    } else {
                        const unsigned long long idx = atomicAdd(&a.ovf_n[t], (unsigned long long)parts);
                        if ((long long)(idx + parts) <= a.ovf_cap) {
                            for (int k = 0; k < parts; k++) a.tier[t][idx + k] = (long long)v0 * CSPLIT_ENC + k;
                        } else {
                            atomicOr(a.err, 4u);
                        }
                    }
                }
                continue;
            }
            const long long c0 = tl ? clock64() : 0;
            const long long rb = a.off[v0];
            const int w = row_words(d);
            for (int k = ln; k < d; k += W) L[k] = a.adj[rb + k];
            for (int k = ln; k < d * w; k += W) rows[k] = 0u;
            if (ACC && ln == 0) bytes += 4ull * d + 16ull;
            wsync();
            vtx_rows<ACC>(a, L, d, rows, w, 0, 1, ln, bytes);
            wsync();
            // L is no longer needed: it holds the pair offsets from here on
            int *off = L;
            int run = 0;
            for (int pb = 0; pb < d; pb += W) {
                const int p = pb + ln;
                const int c = p < d ? row_popc(rows, p, w) : 0;
                if (ACC && p < d) bytes += 4ull * ((p >> 5) + 1);
                int tot = 0;
                const int x = excl_scan(c, &tot);
                if (p < d) off[p] = run + x;
                run += tot;
            }
            if (ln == 0) {
                off[d] = run;
                acc[3] += (unsigned long long)run;
            }
            wsync();
            const long long c1 = tl ? clock64() : 0;
            if (a.kmax >= 4)
                vtx_pairs<row_words(32 * CMAXW), ACC>(a, (const unsigned *)rows, d, w, off, run, ln, W, acc, bytes);
            wsync();
            if (tl) {
                tl_build += c1 - c0;
                tl_count += clock64() - c1;
                tl_items++;
                tl_pairs += run;
            }
        }
    }
    clique_flush<ACC>(a, acc, bytes, ln);
    if (tl && ln == 0) {
        const int unit = blockIdx.x * (blockDim.x / W) + wid;
        tl_write(a.tl, a.tl_base + unit, 0, unit, tl_t0, wall_clock64(), tl_build, tl_count, tl_items, tl_pairs);
    }
}

// Workgroup tiers: the NW waves of a workgroup share one vertex of tier
// list TIER. Rows MW words wide (d <= 32 * MW) sit in LDS or, without LDS,
// in the workgroup's HBM scratch; MW == 0 (any d up to hcap) reads every
// word from HBM as needed. The waves split the rows between them for the
// build and take pairs round-robin for the count.
template <int MW, bool LDS, int TIER, int NW, int MODE>
__global__ void __launch_bounds__(64 * NW, NW == CBLOCKW0 ? 8 : 1) clique_block_kernel(CArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / W, ln = threadIdx.x % W;
    constexpr int lcap = 32 * MW;
    // LDS: the list (then the pair offsets), then the rows
    int *Ls = reinterpret_cast<int *>(smem);
    unsigned *rows_l = reinterpret_cast<unsigned *>(Ls + (LDS ? lcap + 4 : 0));
    // the index of L after the rows (LDS tiers): one entry per thread
    static_assert(!LDS || lcap <= 64 * NW, "one index entry per thread");
    unsigned short *ix = LDS ? reinterpret_cast<unsigned short *>(rows_l + lcap * row_words(lcap)) : nullptr;
    constexpr unsigned ix_mask = (unsigned)index_slots(lcap) - 1;
    static_assert((index_slots(lcap) & (index_slots(lcap) - 1)) == 0, "index slots are a power of two");
    __shared__ long long vsh;
    __shared__ int wtot[NW];
    const int hws = row_words(a.hcap);
    unsigned *rows_g = LDS ? nullptr : a.grows + (long long)blockIdx.x * a.hcap * hws;
    int *off_g = LDS ? nullptr : a.goff + (long long)blockIdx.x * (a.hcap + 1);
    unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0}, bytes = 0;
    constexpr bool ACC = MODE & 1, tl = (MODE & 2) != 0;
    const long long tl_t0 = tl ? wall_clock64() : 0;
    long long tl_build = 0, tl_count = 0, tl_items = 0, tl_pairs = 0, c0 = 0, c1 = 0;
    for (;;) {
        if (tl) c0 = clock64();
        if (threadIdx.x == 0) {
            const unsigned long long r = atomicAdd(&a.queue[1 + TIER], 1ull);
            const unsigned long long cnt = a.ovf_n[TIER];
            vsh = (r < cnt && (long long)r < a.ovf_cap) ? a.tier[TIER][r] : -1;
        }
        __syncthreads();
        if (vsh < 0) break;
        const int v0 = (int)(vsh / CSPLIT_ENC);
        const int part = (int)(vsh % CSPLIT_ENC);
        const int d = a.split[v0];
        const int parts = clique_parts(TIER, d);
        const long long rb = a.off[v0];
        const int w = row_words(d);
        if (d > (MW > 0 ? lcap : a.hcap) || d > a.hcap) {  // cannot happen: lists are split by d
            if (threadIdx.x == 0) atomicOr(a.err, 8u);
        } else {
            if (LDS)
                for (int k = threadIdx.x; k < d; k += 64 * NW) Ls[k] = a.adj[rb + k];
            unsigned *rows = LDS ? rows_l : rows_g;
            for (long long k = threadIdx.x; k < (long long)d * w; k += 64 * NW) rows[k] = 0u;
            if (LDS)
                for (int k = threadIdx.x; k < index_slots(lcap); k += 64 * NW) ix[k] = 0;
            // every part builds the rows; W_alg charges one build
            unsigned long long build_bytes = 0;
            if (ACC && threadIdx.x == 0) build_bytes += 4ull * d + 16ull;
            if (!LDS) __threadfence();
            __syncthreads();
            if (LDS) {
                // thread i enters L[i]: each round it writes 1 + i to its slot
                // if the slot is empty, and after a barrier moves on if
                // another thread's entry holds the slot. Entries are never
                // overwritten, and one moves past a slot only when that slot
                // is taken, so every slot between an entry's hash and its
                // place stays full, as linear probing needs.
                const int i = threadIdx.x;
                bool pending = i < d;
                unsigned h = pending ? index_hash(Ls[i], ix_mask) : 0u;
                while (__syncthreads_or(pending)) {
                    if (pending && ix[h] == 0) ix[h] = (unsigned short)(i + 1);
                    __syncthreads();
                    if (pending) {
                        if (ix[h] == i + 1) pending = false;
                        else h = (h + 1) & ix_mask;
                    }
                }
            }
            const int *L = LDS ? Ls : a.adj + rb;
            vtx_rows<ACC>(a, L, d, rows, w, wid, NW, ln, build_bytes, ix, ix_mask);
            if (!LDS) __threadfence();
            __syncthreads();
            // pair offsets: block-wide exclusive prefix of popc(M_p)
            int *off = LDS ? Ls : off_g;
            int run = 0;
            for (int pb = 0; pb < d; pb += 64 * NW) {
                const int p = pb + (int)threadIdx.x;
                const int c = p < d ? row_popc((const unsigned *)rows, p, w) : 0;
                if (ACC && p < d) build_bytes += 4ull * ((p >> 5) + 1);
                int tot = 0;
                const int x = excl_scan(c, &tot);
                if (ln == 0) wtot[wid] = tot;
                __syncthreads();
                int pre = 0, all = 0;
                for (int k = 0; k < NW; k++) {
                    const int t = wtot[k];
                    if (k < wid) pre += t;
                    all += t;
                }
                if (p < d) off[p] = run + pre + x;
                run += all;
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                off[d] = run;
                if (part == 0) acc[3] += (unsigned long long)run;
            }
            if (part == 0) bytes += build_bytes;
            if (!LDS) __threadfence();
            __syncthreads();
            if (tl) c1 = clock64();
            if (a.kmax >= 4)
                vtx_pairs<(MW > 0 ? row_words(32 * MW) : 0), ACC>(a, (const unsigned *)rows, d, w, off, run,
                                                           part * 64 * NW + (int)threadIdx.x,
                          64 * NW * parts, acc, bytes);
            if (tl) {
                tl_pairs += run;
                tl_items++;
            }
        }
        __syncthreads();
        if (tl && c1) {
            tl_build += c1 - c0;
            tl_count += clock64() - c1;
            c1 = 0;
        }
    }
    clique_flush<ACC>(a, acc, bytes, ln);
    if (tl && threadIdx.x == 0)
        tl_write(a.tl, a.tl_base + blockIdx.x, 1 + TIER, blockIdx.x, tl_t0, wall_clock64(), tl_build, tl_count,
                 tl_items, tl_pairs);
}

__global__ void fill_u64(unsigned long long *p, long long n, unsigned long long v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void split_kernel(const long long *off, const int *adj, int n, int *split) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        split[v] = lbound(adj + off[v], (int)(off[v + 1] - off[v]), (int)v);
}

// Bitmap row h of hub vertex hub_vertex[h] (or h when hub_vertex is null);
// rows start zeroed.
__global__ void hub_rows_kernel(const long long *off, const int *adj, const int *hub_vertex, int n_hubs,
                                unsigned *bits, long long words) {
    for (int h = blockIdx.x; h < n_hubs; h += gridDim.x) {
        const int v = hub_vertex ? hub_vertex[h] : h;
        unsigned *row = bits + (long long)h * words;
        for (long long k = off[v] + threadIdx.x; k < off[v + 1]; k += blockDim.x) {
            const int x = adj[k];
            atomicOr(&row[x >> 5], 1u << (x & 31));
        }
    }
}

__global__ void copy_kernel(const int4 *__restrict__ src, int4 *__restrict__ dst, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace gatk

// ===========================================================================
// Host side
// ===========================================================================
namespace gat {

#define HIPCHECK(x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) GAT_FAIL(GAT_ENODEV, "%s failed: %s", #x, hipGetErrorString(e_));      \
    } while (0)

struct DeviceInfo {
    bool ok = false;
    int cus = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
    long long wall_khz = 0;  // wall_clock64() ticks per millisecond
    char arch[64] = {};
};

static int device_info(DeviceInfo *out) {
    static DeviceInfo info;
    static std::once_flag once;
    static int rc = GAT_OK;
    std::call_once(once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            set_error("no HIP device visible");
            rc = GAT_ENODEV;
            return;
        }
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) != hipSuccess) {
            set_error("hipGetDeviceProperties failed");
            rc = GAT_ENODEV;
            return;
        }
        info.cus = p.multiProcessorCount;
        info.lds_per_block = p.sharedMemPerBlock;
        info.lds_per_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : p.sharedMemPerBlock;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || khz <= 0) khz = 100000;
        info.wall_khz = khz;
        std::snprintf(info.arch, sizeof(info.arch), "%s", p.gcnArchName);
        if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
            set_error("libgat is built for gfx950; device is %s", p.gcnArchName);
            rc = GAT_ENODEV;
            return;
        }
        info.ok = true;
    });
    if (rc) {
        set_error("no usable gfx950 device");
        return rc;
    }
    *out = info;
    return GAT_OK;
}

struct DeviceGraph {
    long long *off = nullptr;
    int *adj = nullptr;
    int *split = nullptr;
    // hub bitmaps, built for the threshold hub_degree (0 = not built)
    int hub_degree = 0;
    int n_hubs = 0;
    long long hub_words = 0;
    unsigned *hub_bits = nullptr;
    int *hub_map = nullptr;     // null when the graph is degree-ordered
    int *hub_vertex = nullptr;
    void free_hubs() {
        for (void *p : {(void *)hub_bits, (void *)hub_map, (void *)hub_vertex})
            if (p) (void)hipFree(p);
        hub_bits = nullptr;
        hub_map = hub_vertex = nullptr;
        n_hubs = 0;
        hub_degree = 0;
    }
    ~DeviceGraph() {
        free_hubs();
        if (off) (void)hipFree(off);
        if (adj) (void)hipFree(adj);
        if (split) (void)hipFree(split);
    }
};

struct PlanDevice {
    hipStream_t stream = nullptr;
    GatPlanDev *plan = nullptr;
    unsigned long long *counts = nullptr, *queue = nullptr, *ovf_n = nullptr, *bytes = nullptr;
    unsigned int *err = nullptr;
    long long *ovf = nullptr;
    gatk::SplitTask *subs = nullptr;  // split subtasks (array store)
    gatk::KArgs *dargs = nullptr;     // arguments of the array-store launches of a count
    long long *tl = nullptr;          // timeline records of the last GAT_FLAG_TIMELINE count
    int tl_n = 0;                     // records that count reserved (0: none)
    long long *tier[3] = {};  // clique vertices of the workgroup tiers
    long long ovf_cap = 0;
    int *gstore = nullptr;
    size_t gstore_bytes = 0;
    int ovf_grid = 0;
    hipEvent_t ev[4] = {};
    // clique HBM-tier rows and pair offsets
    unsigned *grows = nullptr;
    int *goff = nullptr;
    int big_waves = 0, big_cap = 0;
    std::shared_ptr<DeviceGraph> dg;
    // captured graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
};

void destroy_plan_device(PlanDevice *d) {
    if (!d) return;
    if (d->gexec) (void)hipGraphExecDestroy(d->gexec);
    if (d->graph) (void)hipGraphDestroy(d->graph);
    for (auto &e : d->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : {(void *)d->plan, (void *)d->dargs, (void *)d->tl, (void *)d->counts, (void *)d->queue, (void *)d->ovf_n, (void *)d->bytes,
                    (void *)d->err, (void *)d->ovf, (void *)d->subs, (void *)d->tier[0], (void *)d->tier[1], (void *)d->tier[2],
                    (void *)d->gstore, (void *)d->grows,
                    (void *)d->goff})
        if (p) (void)hipFree(p);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

static int ensure_device_graph(Graph &g, std::shared_ptr<DeviceGraph> *out) {
    if (g.dev) {
        *out = std::static_pointer_cast<DeviceGraph>(g.dev);
        return GAT_OK;
    }
    auto dg = std::make_shared<DeviceGraph>();
    HIPCHECK(hipMalloc(&dg->off, sizeof(long long) * (size_t(g.n) + 1)));
    HIPCHECK(hipMalloc(&dg->adj, sizeof(int) * std::max<size_t>(1, size_t(g.m))));
    HIPCHECK(hipMalloc(&dg->split, sizeof(int) * std::max<size_t>(1, size_t(g.n))));
    HIPCHECK(hipMemcpy(dg->off, g.off.data(), sizeof(long long) * (size_t(g.n) + 1), hipMemcpyHostToDevice));
    if (g.m) HIPCHECK(hipMemcpy(dg->adj, g.adj.data(), sizeof(int) * size_t(g.m), hipMemcpyHostToDevice));
    if (g.n) {
        hipLaunchKernelGGL(gatk::split_kernel, dim3(1024), dim3(256), 0, nullptr, dg->off, dg->adj, g.n, dg->split);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipDeviceSynchronize());
    g.dev = dg;
    *out = dg;
    return GAT_OK;
}

// Neighbourhood bitmaps for every vertex of degree >= thr, most connected
// first, within kHubBytes. Built once per (graph, threshold) outside the
// timed count, like the graph upload itself.
constexpr double kHubBytes = 8.0 * (1ull << 30);

static int ensure_hubs(const Graph &g, DeviceGraph *dg, int thr) {
    if (dg->hub_degree == thr) return GAT_OK;
    dg->free_hubs();
    const long long words = ((long long)g.n + 31) / 32;
    bool ordered = true;  // degree-ordered graphs keep hubs as a prefix
    for (int32_t v = 1; v < g.n && ordered; v++) ordered = g.deg(v) <= g.deg(v - 1);
    std::vector<int32_t> hubs;
    if (ordered) {
        int32_t h = 0;
        while (h < g.n && g.deg(h) >= thr) h++;
        hubs.resize(h);
        std::iota(hubs.begin(), hubs.end(), 0);
    } else {
        for (int32_t v = 0; v < g.n; v++)
            if (g.deg(v) >= thr) hubs.push_back(v);
        std::stable_sort(hubs.begin(), hubs.end(), [&](int32_t x, int32_t y) { return g.deg(x) > g.deg(y); });
    }
    const long long max_hubs = (long long)(kHubBytes / (4.0 * std::max(1ll, words)));
    if ((long long)hubs.size() > max_hubs) hubs.resize(max_hubs);
    dg->hub_degree = thr;
    dg->hub_words = words;
    if (hubs.empty()) return GAT_OK;
    const long long n_u32 = (long long)hubs.size() * words;
    HIPCHECK(hipMalloc(&dg->hub_bits, sizeof(unsigned) * (size_t)((n_u32 + 1) / 2 * 2)));
    hipLaunchKernelGGL(gatk::fill_u64, dim3(2048), dim3(256), 0, nullptr, (unsigned long long *)dg->hub_bits,
                       (n_u32 + 1) / 2, 0ull);
    if (!ordered) {
        std::vector<int32_t> map(size_t(g.n), -1);
        for (size_t h = 0; h < hubs.size(); h++) map[hubs[h]] = int32_t(h);
        HIPCHECK(hipMalloc(&dg->hub_map, sizeof(int) * size_t(g.n)));
        HIPCHECK(hipMalloc(&dg->hub_vertex, sizeof(int) * hubs.size()));
        HIPCHECK(hipMemcpy(dg->hub_map, map.data(), sizeof(int) * size_t(g.n), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dg->hub_vertex, hubs.data(), sizeof(int) * hubs.size(), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(gatk::hub_rows_kernel, dim3((unsigned)std::min<size_t>(hubs.size(), 65535)), dim3(256), 0,
                       nullptr, dg->off, dg->adj, dg->hub_vertex, (int)hubs.size(), dg->hub_bits, words);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
    dg->n_hubs = (int)hubs.size();
    return GAT_OK;
}

// Split subtasks: capacity of each level's list, and the default time a
// task runs in the tail of a launch before it hands on its unfinished loops.
// Subtask counters live at ovf_n[8 + k] (k = 1 .. levels: LDS levels, then
// the HBM list).
constexpr long long kSubCap = 1ll << 20;
constexpr int kLevels = gatk::kSplitLevels;
constexpr int kSplitUs = 2000;
// device argument records of one array-store count: the main launch, the
// LDS subtask levels, the HBM launch and the HBM subtask launch
constexpr int kArgSlots = kLevels + 3;
// timeline capacity (records of gatk::kTlFields int64 each)
constexpr int kTlCap = 1 << 15;

struct Launch {
    gatk::KArgs a;
    dim3 grid, block;
    size_t lds = 0;
    int wave_bytes = 0;
    dim3 ovf_grid;
    size_t ovf_lds = 0;
    int ovf_wave_bytes = 0;
    gatk::KArgs ao;
    bool clique = false;  // bitmap clique path
    gatk::CArgs ca;
    size_t clique_lds = 0;
    int clique_wave_bytes = 0;
    dim3 mid_grid;      // workgroup tier 0 (rows in LDS)
    size_t mid_lds = 0;
    dim3 t1_grid;       // workgroup tier 1 (LDS when t1_lds > 0, else HBM)
    size_t t1_lds = 0;
    dim3 big_grid;      // workgroup tier 2 (HBM)
    long long tasks = 0;  // root edges of the count
};

// Restrict a prepared count to the root vertices [lo, hi): the array store
// takes the directed edges of those vertices, the bitmap store those top
// vertices.
static void set_roots(Launch *L, const Graph &g, int32_t lo, int32_t hi) {
    L->a.e_lo = L->ao.e_lo = g.off[lo];
    L->a.e_hi = L->ao.e_hi = g.off[hi];
    L->ca.root_lo = lo;
    L->ca.root_hi = hi;
    L->tasks = g.off[hi] - g.off[lo];
}

static void fill_defaults(gat_kernel_config *c) {
    if (c->block_waves <= 0) c->block_waves = 1;
    if (c->waves_per_cu <= 0) c->waves_per_cu = 8;
    if (c->queue_chunk <= 0) c->queue_chunk = 16;
    if (c->slot_capacity <= 0) c->slot_capacity = 1024;
    if (c->intersect_strategy < 0 || c->intersect_strategy > 4) c->intersect_strategy = 0;
    if (c->gallop_ratio <= 0) c->gallop_ratio = 32;
    if (c->search_linear_threshold < 0) c->search_linear_threshold = 0;
    if (c->hub_degree == 0) c->hub_degree = 256;
    if (c->split_us == 0) c->split_us = kSplitUs;
}

// Knobs and launch shapes of the clique store: the main launch (one wave
// per vertex, rows in the wave's LDS slice) and workgroup tiers 0 and 1
// (rows in LDS; t1_lds = 0 when the device cannot give tier 1 its LDS).
static int clique_shape(const DeviceInfo &di, const gat_kernel_config &cfg, Launch *L) {
    gatk::CArgs &c = L->ca;
    c.cap = std::max(32, std::min(32 * gatk::CMAXW, cfg.slot_capacity / 32 * 32));
    c.chunk = cfg.queue_chunk;
    c.lin_thr = cfg.search_linear_threshold;
    c.gallop_ratio = cfg.gallop_ratio;
    L->block = dim3(64 * cfg.block_waves);
    L->grid = dim3(std::max(1, std::max(1, cfg.waves_per_cu * di.cus) / cfg.block_waves));
    L->clique_wave_bytes = (4 * (c.cap + 4) + 4 * c.cap * gatk::row_words(c.cap) + 15) / 16 * 16;
    L->clique_lds = size_t(L->clique_wave_bytes) * cfg.block_waves;
    if (L->clique_lds > di.lds_per_block)
        GAT_FAIL(GAT_EINFEASIBLE, "bitmap config needs %zu B of LDS per workgroup", L->clique_lds);
    L->mid_lds = (size_t)4 * (32 * gatk::CMIDW + 4) + (size_t)4 * 32 * gatk::CMIDW * gatk::row_words(32 * gatk::CMIDW) +
                 (size_t)2 * gatk::index_slots(32 * gatk::CMIDW);
    if (L->mid_lds > di.lds_per_block) GAT_FAIL(GAT_EINFEASIBLE, "device LDS too small for the clique block tier");
    L->mid_grid = dim3(std::max(1, (int)(di.lds_per_cu / L->mid_lds)) * di.cus);
    const size_t t1 = (size_t)4 * (32 * gatk::CBIGW + 4) + (size_t)4 * 32 * gatk::CBIGW * gatk::row_words(32 * gatk::CBIGW) +
                      (size_t)2 * gatk::index_slots(32 * gatk::CBIGW);
    L->t1_lds = t1 <= di.lds_per_block ? t1 : 0;
    if (L->t1_lds) L->t1_grid = dim3(std::max(1, (int)(di.lds_per_cu / t1)) * di.cus);
    return GAT_OK;
}

static int prepare(gat_plan *p, const gat_kernel_config *cfg_in, Launch *L) {
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc) return rc;
    gat_kernel_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    // the array store's waves wait on dependent list reads: its default grid
    // keeps twice the clique store's waves per CU (16 against 8)
    if (cfg.waves_per_cu <= 0 && (p->cp.clique_ks == 0 || cfg.vertex_store == 1)) cfg.waves_per_cu = 16;
    fill_defaults(&cfg);
    if (cfg.block_waves > 4) GAT_FAIL(GAT_EINFEASIBLE, "block_waves %d > 4", cfg.block_waves);
    Graph &g = *p->graph;
    const GatPlanDev &pd = p->cp.dev;
    if (!p->dev) {
        auto *d = new PlanDevice();
        p->dev = d;
        rc = ensure_device_graph(g, &d->dg);
        if (rc) return rc;
        HIPCHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        HIPCHECK(hipMalloc(&d->plan, sizeof(GatPlanDev)));
        HIPCHECK(hipMemcpy(d->plan, &pd, sizeof(GatPlanDev), hipMemcpyHostToDevice));
        HIPCHECK(hipMalloc(&d->counts, sizeof(unsigned long long) * GATP_MAX_TERMS));
        HIPCHECK(hipMalloc(&d->queue, sizeof(unsigned long long) * 16));
        HIPCHECK(hipMalloc(&d->ovf_n, sizeof(unsigned long long) * 16));
        HIPCHECK(hipMalloc(&d->bytes, sizeof(unsigned long long) * 4));
        HIPCHECK(hipMalloc(&d->err, sizeof(unsigned int) * 2));
        HIPCHECK(hipMalloc(&d->dargs, sizeof(gatk::KArgs) * kArgSlots));
        for (auto &e : d->ev) HIPCHECK(hipEventCreate(&e));
    }
    PlanDevice *d = p->dev;
    const long long chunk = cfg.queue_chunk;
    const bool clique = p->cp.clique_ks != 0 && cfg.vertex_store != 1;
    if (!clique && !d->gstore) {
        // HBM set buffers of the overflow launch, where every set fits in the
        // maximum degree: 16 waves per CU hide the latency of HBM-resident
        // sets, within 4 GiB of buffers.
        const size_t per_wave = (size_t)std::max(1, pd.n_store) * (size_t)std::max(1, g.max_deg) * 4;
        int waves = std::max(64, 16 * di.cus);
        while (waves > 64 && (double)waves * per_wave > 4.0 * (1 << 30)) waves /= 2;
        d->ovf_grid = waves;
        d->gstore_bytes = (size_t)waves * per_wave;
        HIPCHECK(hipMalloc(&d->gstore, std::max<size_t>(d->gstore_bytes, 64)));
    }
    if (cfg.vertex_store == 2 && !clique)
        GAT_FAIL(GAT_EINFEASIBLE, "bitmap vertex store needs a plan of cliques K3..K6 in unique or labeled mode");
    // single-child roots hand over single edges (run_task), others whole tasks
    const bool edge_grain = pd.nodes[0].child_count == 1;
    // overflow entries: edges or tasks, and split subtasks
    const long long ovf_cap = clique ? 1 : (edge_grain ? g.m + 1 : (g.m + chunk - 1) / chunk + 1);
    const bool split = !clique && cfg.split_us > 0;
    // subtask lists: the LDS levels (kSubCap each), then the HBM list (LDS
    // overflow of the levels plus the HBM launch's splits, 3 kSubCap)
    if (split && !d->subs) HIPCHECK(hipMalloc(&d->subs, sizeof(gatk::SplitTask) * (kLevels + 3) * kSubCap));
    if (ovf_cap > d->ovf_cap) {
        if (d->ovf) HIPCHECK(hipFree(d->ovf));
        for (auto &t : d->tier)
            if (t) HIPCHECK(hipFree(t));
        d->ovf = nullptr;
        for (auto &t : d->tier) t = nullptr;
        HIPCHECK(hipMalloc(&d->ovf, sizeof(long long) * ovf_cap));
        d->ovf_cap = ovf_cap;
    }
    if (clique && !d->tier[0])
        for (auto &t : d->tier) HIPCHECK(hipMalloc(&t, sizeof(long long) * (size_t)(g.n + 1) * gatk::CSPLIT_ENC));
    const int wave_bytes = int((sizeof(gatk::WaveState) + size_t(pd.n_store) * cfg.slot_capacity * 4 + 15) / 16 * 16);
    const size_t lds = size_t(wave_bytes) * cfg.block_waves;
    if (lds > di.lds_per_block)
        GAT_FAIL(GAT_EINFEASIBLE, "config needs %zu B of LDS per workgroup (device allows %zu)", lds,
                 di.lds_per_block);
    gatk::KArgs a{};
    a.off = d->dg->off;
    a.adj = d->dg->adj;
    a.split = d->dg->split;
    a.n = g.n;
    a.m = g.m;
    a.e_lo = 0;
    a.e_hi = g.m;
    a.plan = d->plan;
    a.counts = d->counts;
    a.queue = d->queue;
    a.ovf = d->ovf;
    a.ovf_n = d->ovf_n;
    a.ovf_cap = d->ovf_cap;
    a.gstore = d->gstore;
    a.bytes = d->bytes;
    a.err = d->err;
    a.cap = cfg.slot_capacity;
    a.chunk = cfg.queue_chunk;
    a.lin_thr = cfg.search_linear_threshold;
    a.gallop_ratio = cfg.gallop_ratio;
    a.strategy = cfg.intersect_strategy;
    a.account = 0;
    a.sub_mode = 0;
    a.subs = d->subs;
    a.in_n = nullptr;
    a.out_n = d->ovf_n + 8 + 1;  // the main launch splits into level 1
    a.in_base = a.out_base = 0;
    a.in_cap = a.out_cap = kSubCap;
    a.sovf_n = d->ovf_n + 8 + kLevels + 1;
    a.sovf_base = kLevels * kSubCap;
    a.sovf_cap = 3 * kSubCap;
    a.split_ticks = split ? (long long)cfg.split_us * di.wall_khz / 1000 : 0;
    if (!clique && cfg.hub_degree > 0) {
        rc = ensure_hubs(g, d->dg.get(), cfg.hub_degree);
        if (rc) return rc;
        a.hub_bits = d->dg->hub_bits;
        a.hub_map = d->dg->hub_map;
        a.hub_words = d->dg->hub_words;
        a.n_hubs = d->dg->n_hubs;
    }
    L->a = a;
    L->block = dim3(64 * cfg.block_waves);
    const int total_waves = std::max(1, cfg.waves_per_cu * di.cus);
    L->grid = dim3(std::max(1, total_waves / cfg.block_waves));
    L->lds = lds;
    L->wave_bytes = wave_bytes;
    L->ao = a;
    L->ao.cap = std::max(1, g.max_deg);
    if (edge_grain) L->ao.chunk = 1;
    L->ovf_grid = dim3(d->ovf_grid);
    L->ovf_wave_bytes = int((sizeof(gatk::WaveState) + 15) / 16 * 16);
    L->ovf_lds = L->ovf_wave_bytes;
    L->clique = clique;
    if (clique) {
        gatk::CArgs c{};
        c.off = a.off;
        c.adj = a.adj;
        c.split = a.split;
        c.n = a.n;
        c.m = a.m;
        c.root_lo = 0;
        c.root_hi = a.n;
        c.counts = a.counts;
        for (int k = 0; k < 8; k++) c.pat_of_k[k] = p->cp.clique_pattern[k];
        c.kmax = 31 - __builtin_clz(p->cp.clique_ks);
        c.queue = a.queue;
        for (int t = 0; t < 3; t++) c.tier[t] = d->tier[t];
        c.ovf_n = a.ovf_n;
        c.ovf_cap = (long long)(g.n + 1) * gatk::CSPLIT_ENC;
        c.bytes = a.bytes;
        c.err = a.err;
        // HBM tier: rows sized by the maximum out-degree
        if (g.max_out < 0) {
            int32_t mo = 0;
            for (int32_t v = 0; v < g.n; v++) {
                const int32_t *b = g.adj.data() + g.off[v], *en = g.adj.data() + g.off[v + 1];
                mo = std::max(mo, int32_t(std::lower_bound(b, en, v) - b));
            }
            g.max_out = mo;
        }
        const int hcap = std::max(32, g.max_out);
        const int ws = gatk::row_words(hcap);
        int waves = std::max(64, 8 * di.cus);
        while (waves > 64 && (double)waves * hcap * ws * 4 > 16e9) waves /= 2;
        if (d->big_cap < hcap || d->big_waves < waves) {
            if (d->grows) HIPCHECK(hipFree(d->grows));
            if (d->goff) HIPCHECK(hipFree(d->goff));
            d->grows = nullptr;
            d->goff = nullptr;
            HIPCHECK(hipMalloc(&d->grows, (size_t)waves * hcap * ws * 4));
            HIPCHECK(hipMalloc(&d->goff, (size_t)waves * (hcap + 1) * 4));
            d->big_cap = hcap;
            d->big_waves = waves;
        }
        c.grows = d->grows;
        c.goff = d->goff;
        c.hcap = d->big_cap;
        L->ca = c;
        rc = clique_shape(di, cfg, L);
        if (rc) return rc;
        L->big_grid = dim3(std::min(d->big_waves, 2 * di.cus));  // one scratch area per workgroup
        if (!L->t1_lds) L->t1_grid = L->big_grid;
    }
    set_roots(L, g, 0, g.n);
    return GAT_OK;
}

// Enqueue the whole count on the plan's stream (capturable: no memsets).
static int enqueue(PlanDevice *d, const Launch &L, bool account, bool events, bool timeline = false) {
    hipStream_t s = d->stream;
    // timeline: one record per wave (array store, clique main launch) or per
    // workgroup (clique tiers), launch after launch
    long long *tl = nullptr;
    int tl_next = 0;
    auto tl_take = [&](int units) {  // first record of the next launch
        const int b = tl_next;
        tl_next += units;
        return b;
    };
    if (timeline) {
        const int waves = (int)(L.grid.x * (L.block.x / 64));
        const long long need = L.clique ? (long long)waves + L.mid_grid.x + L.t1_grid.x + L.big_grid.x
                                        : (long long)waves * (1 + (L.a.split_ticks > 0 ? kLevels : 0)) +
                                              (long long)L.ovf_grid.x * (L.a.split_ticks > 0 ? 2 : 1);
        if (need > kTlCap) GAT_FAIL(GAT_ELIMIT, "timeline needs %lld records (> %d)", need, kTlCap);
        if (!d->tl) HIPCHECK(hipMalloc(&d->tl, sizeof(long long) * gatk::kTlFields * kTlCap));
        hipLaunchKernelGGL(gatk::fill_u64, dim3(64), dim3(256), 0, s, (unsigned long long *)d->tl,
                           (long long)gatk::kTlFields * kTlCap, ~0ull);
        tl = d->tl;
    }
    hipLaunchKernelGGL(gatk::fill_u64, dim3(1), dim3(64), 0, s, d->counts, (long long)GATP_MAX_TERMS, 0ull);
    hipLaunchKernelGGL(gatk::fill_u64, dim3(1), dim3(64), 0, s, d->queue, 16ll, 0ull);
    hipLaunchKernelGGL(gatk::fill_u64, dim3(1), dim3(64), 0, s, d->ovf_n, 16ll, 0ull);
    hipLaunchKernelGGL(gatk::fill_u64, dim3(1), dim3(64), 0, s, d->bytes, 4ll, 0ull);
    hipLaunchKernelGGL(gatk::fill_u64, dim3(1), dim3(64), 0, s, (unsigned long long *)d->err, 1ll, 0ull);
    gatk::KArgs a = L.a, ao = L.ao;
    a.account = ao.account = account ? 1 : 0;
    a.tl = ao.tl = tl;
    ao.bytes = d->bytes + 1;  // second launch: its own W_alg counter
    const int waves = (int)(L.grid.x * (L.block.x / 64));
    if (events) HIPCHECK(hipEventRecord(d->ev[0], s));
    gatk::CArgs c = L.ca;
    c.account = account ? 1 : 0;
    c.tl = tl;
    if (L.clique) {
        // W_alg accounting and timeline records come from separate kernel
        // instantiations (MODE bits 0 and 1)
        auto run = [&](auto mode_tag) -> int {
            constexpr int T = decltype(mode_tag)::value;
            c.tl_base = tl_take(waves);
            hipLaunchKernelGGL(gatk::clique_kernel<T>, L.grid, L.block, L.clique_lds, s, c, L.clique_wave_bytes);
            if (events) HIPCHECK(hipEventRecord(d->ev[1], s));
            c.bytes = d->bytes + 1;  // workgroup tier 0: its own W_alg counter
            c.tl_base = tl_take((int)L.mid_grid.x);
            hipLaunchKernelGGL((gatk::clique_block_kernel<gatk::CMIDW, true, 0, gatk::CBLOCKW0, T>), L.mid_grid,
                               dim3(64 * gatk::CBLOCKW0), L.mid_lds, s, c);
            if (events) HIPCHECK(hipEventRecord(d->ev[3], s));
            c.bytes = d->bytes + 2;  // tiers 1 and 2
            c.tl_base = tl_take((int)L.t1_grid.x);
            if (L.t1_lds)
                hipLaunchKernelGGL((gatk::clique_block_kernel<gatk::CBIGW, true, 1, gatk::CBLOCKW1, T>), L.t1_grid,
                                   dim3(64 * gatk::CBLOCKW1), L.t1_lds, s, c);
            else
                hipLaunchKernelGGL((gatk::clique_block_kernel<gatk::CBIGW, false, 1, gatk::CBLOCKW1, T>), L.t1_grid,
                                   dim3(64 * gatk::CBLOCKW1), 0, s, c);
            c.tl_base = tl_take((int)L.big_grid.x);
            hipLaunchKernelGGL((gatk::clique_block_kernel<0, false, 2, gatk::CBLOCKW1, T>), L.big_grid,
                               dim3(64 * gatk::CBLOCKW1), 0, s, c);
            return GAT_OK;
        };
        const int mode = (account ? 1 : 0) | (timeline ? 2 : 0);
        const int rc = mode == 0   ? run(std::integral_constant<int, 0>{})
                       : mode == 1 ? run(std::integral_constant<int, 1>{})
                       : mode == 2 ? run(std::integral_constant<int, 2>{})
                                   : run(std::integral_constant<int, 3>{});
        if (rc) return rc;
    } else {
        a.tl_base = tl_take(waves);
        hipLaunchKernelGGL(gatk::put_kargs, dim3(1), dim3(64), 0, s, a, d->dargs);
        hipLaunchKernelGGL(gatk::engine_kernel<false>, L.grid, L.block, L.lds, s, d->dargs, L.wave_bytes);
        if (events) HIPCHECK(hipEventRecord(d->ev[1], s));
        if (a.split_ticks > 0)  // split subtasks, level by level, in LDS like the main launch
            for (int k = 1; k <= kLevels; k++) {
                gatk::KArgs sk = a;
                sk.sub_mode = k;
                sk.in_n = d->ovf_n + 8 + k;
                sk.in_base = (k - 1) * kSubCap;
                sk.out_n = k < kLevels ? d->ovf_n + 8 + k + 1 : nullptr;  // the last level does not split
                sk.out_base = k * kSubCap;
                sk.bytes = d->bytes + 1;
                sk.tl_base = tl_take(waves);
                hipLaunchKernelGGL(gatk::put_kargs, dim3(1), dim3(64), 0, s, sk, d->dargs + k);
                hipLaunchKernelGGL(gatk::engine_sub_kernel<false>, L.grid, L.block, L.lds, s, d->dargs + k,
                                   L.wave_bytes);
            }
        ao.out_n = d->ovf_n + 8 + kLevels + 1;  // the HBM launch splits into the HBM list
        ao.out_base = kLevels * kSubCap;
        ao.out_cap = 3 * kSubCap;
        ao.tl_base = tl_take((int)L.ovf_grid.x);
        hipLaunchKernelGGL(gatk::put_kargs, dim3(1), dim3(64), 0, s, ao, d->dargs + kLevels + 1);
        hipLaunchKernelGGL(gatk::engine_kernel<true>, L.ovf_grid, dim3(64), L.ovf_lds, s, d->dargs + kLevels + 1,
                           L.ovf_wave_bytes);
        if (ao.split_ticks > 0) {
            gatk::KArgs sh = ao;
            sh.sub_mode = kLevels + 1;  // HBM buffers, no further splitting
            sh.in_n = d->ovf_n + 8 + kLevels + 1;
            sh.in_base = kLevels * kSubCap;
            sh.in_cap = 3 * kSubCap;
            sh.out_n = nullptr;
            sh.tl_base = tl_take((int)L.ovf_grid.x);
            hipLaunchKernelGGL(gatk::put_kargs, dim3(1), dim3(64), 0, s, sh, d->dargs + kLevels + 2);
            hipLaunchKernelGGL(gatk::engine_sub_kernel<true>, L.ovf_grid, dim3(64), L.ovf_lds, s,
                               d->dargs + kLevels + 2, L.ovf_wave_bytes);
        }
        if (events) HIPCHECK(hipEventRecord(d->ev[3], s));
    }
    if (events) HIPCHECK(hipEventRecord(d->ev[2], s));
    HIPCHECK(hipGetLastError());
    d->tl_n = timeline ? tl_next : 0;
    return GAT_OK;
}

// counts_out: the requested counts, or with raw set the engine's raw sums
// (a shard's share of divisor * count, before the divisor and multiplier).
static int finish(gat_plan *p, const Launch &L, uint64_t *counts_out, gat_count_stats *st, bool timed,
                  bool raw_sums = false) {
    PlanDevice *d = p->dev;
    const int np = p->cp.dev.n_patterns;
    unsigned long long raw[GATP_MAX_TERMS], ovf[16] = {}, bytes[3] = {0, 0, 0};
    unsigned int err = 0;
    HIPCHECK(hipMemcpyAsync(raw, d->counts, sizeof(unsigned long long) * np, hipMemcpyDeviceToHost, d->stream));
    HIPCHECK(hipMemcpyAsync(ovf, d->ovf_n, sizeof(ovf), hipMemcpyDeviceToHost, d->stream));
    HIPCHECK(hipMemcpyAsync(bytes, d->bytes, sizeof(bytes), hipMemcpyDeviceToHost, d->stream));
    HIPCHECK(hipMemcpyAsync(&err, d->err, sizeof(err), hipMemcpyDeviceToHost, d->stream));
    HIPCHECK(hipStreamSynchronize(d->stream));
    // flags 2 and 4: a device work list (overflow tasks, subtasks, clique
    // tiers) filled up; 1 and 8: states the host's sizing rules out
    if (err & 6u) GAT_FAIL(GAT_ELIMIT, "a device work list overflowed (engine flags 0x%x)", err);
    if (err) GAT_FAIL(GAT_EINTERNAL, "engine reported an internal inconsistency (flags 0x%x)", err);
    for (int i = 0; i < np && raw_sums; i++)
        if (counts_out) counts_out[i] = raw[i];
    for (int i = 0; i < np && !raw_sums; i++) {
        const PatternInfo &pi = p->cp.patterns[i];
        if (raw[i] % pi.divisor)
            GAT_FAIL(GAT_EINVAL, "internal: raw count %llu of pattern %d not divisible by %llu", raw[i], i,
                     (unsigned long long)pi.divisor);
        if (counts_out) counts_out[i] = raw[i] / pi.divisor * pi.multiplier;
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        if (timed) {
            float ms = 0, ms_main = 0, ms_second = 0;
            HIPCHECK(hipEventElapsedTime(&ms, d->ev[0], d->ev[2]));
            HIPCHECK(hipEventElapsedTime(&ms_main, d->ev[0], d->ev[1]));
            HIPCHECK(hipEventElapsedTime(&ms_second, d->ev[1], d->ev[3]));
            st->seconds = ms / 1000.0;
            st->main_kernel_ms = ms_main;
            st->second_kernel_ms = ms_second;
        }
        st->overflow_tasks = L.clique ? ovf[3] : ovf[0];
        unsigned long long made = 0;
        for (int k = 1; k <= kLevels; k++) made += std::min<unsigned long long>(ovf[8 + k], (unsigned long long)kSubCap);
        made += std::min<unsigned long long>(ovf[8 + kLevels + 1], (unsigned long long)(3 * kSubCap));
        st->split_tasks = L.clique ? 0 : (int32_t)(made - ovf[2]);
        st->split_overflow = L.clique ? 0 : (int32_t)ovf[2];
        for (int k = 0; k < 4; k++)
            st->split_level[k] = L.clique ? 0
                                          : (int64_t)std::min<unsigned long long>(
                                                ovf[8 + 1 + k], (unsigned long long)(k < kLevels ? kSubCap : 3 * kSubCap));
        st->alg_bytes = bytes[0] + bytes[1] + bytes[2];
        st->alg_bytes_main = bytes[0];
        st->alg_bytes_second = bytes[1];
        st->tasks = (uint64_t)L.tasks;
        st->grid_blocks = (int)L.grid.x;
        st->block_threads = (int)L.block.x;
        st->lds_bytes = (int)L.lds;
    }
    return GAT_OK;
}

}  // namespace gat

using namespace gat;

extern "C" int gat_device_count(int32_t *out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (out) *out = n;
    return GAT_OK;
}

extern "C" int gat_count(gat_plan *p, const gat_kernel_config *cfg, uint32_t flags, uint64_t *counts_out,
                         gat_count_stats *stats) {
    if (!p) GAT_FAIL(GAT_EINVAL, "null plan");
    Launch L;
    int rc = prepare(p, cfg, &L);
    if (rc) return rc;
    rc = enqueue(p->dev, L, (flags & GAT_FLAG_ACCOUNT) != 0, true, (flags & GAT_FLAG_TIMELINE) != 0);
    if (rc) return rc;
    return finish(p, L, counts_out, stats, true);
}

extern "C" int gat_plan_timeline(const gat_plan *p, int64_t *out, int64_t cap, int64_t *n_out,
                                 double *ticks_per_ms) {
    if (!p || !n_out) GAT_FAIL(GAT_EINVAL, "null plan or n_out");
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc) return rc;
    PlanDevice *d = p->dev;
    *n_out = 0;
    if (ticks_per_ms) *ticks_per_ms = (double)di.wall_khz;
    if (!d || !d->tl || d->tl_n == 0) return GAT_OK;
    std::vector<long long> rec((size_t)d->tl_n * gatk::kTlFields);
    HIPCHECK(hipStreamSynchronize(d->stream));
    HIPCHECK(hipMemcpy(rec.data(), d->tl, rec.size() * sizeof(long long), hipMemcpyDeviceToHost));
    long long t_min = LLONG_MAX;
    for (int i = 0; i < d->tl_n; i++)
        if (rec[(size_t)i * gatk::kTlFields] >= 0) t_min = std::min(t_min, rec[(size_t)i * gatk::kTlFields + 2]);
    int64_t n = 0;
    for (int i = 0; i < d->tl_n; i++) {
        const long long *r = &rec[(size_t)i * gatk::kTlFields];
        if (r[0] < 0) continue;  // a unit that wrote no record
        if (out && n < cap)
            for (int f = 0; f < gatk::kTlFields; f++) out[n * gatk::kTlFields + f] = (f == 2 || f == 3) ? r[f] - t_min : r[f];
        n++;
    }
    *n_out = n;
    return GAT_OK;
}

extern "C" int gat_count_shard(gat_plan *p, const gat_kernel_config *cfg, uint32_t flags, int32_t root_begin,
                               int32_t root_end, uint64_t *raw_out, gat_count_stats *stats) {
    if (!p) GAT_FAIL(GAT_EINVAL, "null plan");
    const int32_t n = p->graph->n;
    if (root_begin < 0 || root_end < root_begin || root_end > n)
        GAT_FAIL(GAT_EINVAL, "root range [%d, %d) is not within [0, %d)", root_begin, root_end, n);
    Launch L;
    int rc = prepare(p, cfg, &L);
    if (rc) return rc;
    set_roots(&L, *p->graph, root_begin, root_end);
    rc = enqueue(p->dev, L, (flags & GAT_FLAG_ACCOUNT) != 0, true, (flags & GAT_FLAG_TIMELINE) != 0);
    if (rc) return rc;
    return finish(p, L, raw_out, stats, true, true);
}

extern "C" int gat_count_shard_edges(gat_plan *p, const gat_kernel_config *cfg, uint32_t flags, int64_t edge_begin,
                                     int64_t edge_end, uint64_t *raw_out, gat_count_stats *stats) {
    if (!p) GAT_FAIL(GAT_EINVAL, "null plan");
    const long long m = p->graph->m;
    if (edge_begin < 0 || edge_end < edge_begin || edge_end > m)
        GAT_FAIL(GAT_EINVAL, "edge range [%lld, %lld) is not within [0, %lld)", (long long)edge_begin,
                 (long long)edge_end, m);
    Launch L;
    int rc = prepare(p, cfg, &L);
    if (rc) return rc;
    if (L.clique)
        GAT_FAIL(GAT_EUNSUPPORTED, "edge shards need the array store (vertex_store = 1): the bitmap clique store "
                                   "shards by top vertex (gat_count_shard)");
    // the array store's queue runs over root edges; every task finds its
    // root by binary search, so any edge range is a shard
    L.a.e_lo = L.ao.e_lo = edge_begin;
    L.a.e_hi = L.ao.e_hi = edge_end;
    L.tasks = edge_end - edge_begin;
    rc = enqueue(p->dev, L, (flags & GAT_FLAG_ACCOUNT) != 0, true, (flags & GAT_FLAG_TIMELINE) != 0);
    if (rc) return rc;
    return finish(p, L, raw_out, stats, true, true);
}

// ---- allocation-free clique count ---------------------------------------
// Scratch of gat_count_cliques_noalloc, one copy per device for the life of
// the process: kNaHead words of counters (counts [0,8), queues [8,24),
// work-list sizes [24,40), W_alg [40,44), error flags at word 44), then the
// split[] array (int32, kNaMaxV entries), then the lists of workgroup tiers 0
// and 1 (int64, kNaTierCap entries each): 32 MiB and 512 B in all.
constexpr int kNaMaxV = 1 << 22;
constexpr long long kNaTierCap = 1 << 20;
constexpr long long kNaHead = 64;
constexpr long long kNaWords = kNaHead + kNaMaxV / 2 + 2 * kNaTierCap;
static __device__ __align__(256) unsigned long long scratch[kNaWords];

// Device address of this device's scratch, looked up once per device (the
// host-side name `scratch` is only a handle for the runtime, never an address).
static int noalloc_scratch(unsigned long long **out) {
    static std::mutex mu;
    static unsigned long long *addr[64] = {};
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) GAT_FAIL(GAT_ENODEV, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            GAT_FAIL(GAT_ENODEV, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        addr[dev] = static_cast<unsigned long long *>(ptr);
    }
    *out = addr[dev];
    return GAT_OK;
}

extern "C" int gat_count_cliques_noalloc(const int64_t *d_offsets, const int32_t *d_adj, int32_t v_cnt, int32_t kmax,
                                         const gat_kernel_config *cfg_in, void *stream, uint64_t *counts_out) {
    if (!d_offsets || !d_adj || !counts_out) GAT_FAIL(GAT_EINVAL, "null graph or counts pointer");
    if (kmax < 3 || kmax > 6) GAT_FAIL(GAT_EINVAL, "kmax %d is not within [3, 6]", kmax);
    if (v_cnt < 0) GAT_FAIL(GAT_EINVAL, "negative v_cnt");
    if (v_cnt > kNaMaxV)
        GAT_FAIL(GAT_ELIMIT, "v_cnt %d exceeds the allocation-free count's limit of %d vertices", v_cnt, kNaMaxV);
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc) return rc;
    gat_kernel_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    fill_defaults(&cfg);
    if (cfg.block_waves > 4) GAT_FAIL(GAT_EINFEASIBLE, "block_waves %d > 4", cfg.block_waves);
    Launch L;
    rc = clique_shape(di, cfg, &L);
    if (rc) return rc;
    if (!L.t1_lds)
        GAT_FAIL(GAT_EINFEASIBLE, "the allocation-free count needs %d KiB of LDS per workgroup",
                 (int)((4 * (32 * gatk::CBIGW + 4) + 4 * 32 * gatk::CBIGW * gatk::row_words(32 * gatk::CBIGW)) / 1024));
    unsigned long long *h = nullptr;
    rc = noalloc_scratch(&h);
    if (rc) return rc;
    gatk::CArgs &c = L.ca;
    c.off = reinterpret_cast<const long long *>(d_offsets);
    c.adj = d_adj;
    int *split = reinterpret_cast<int *>(h + kNaHead);
    c.split = split;
    c.n = v_cnt;
    c.m = 0;
    c.root_lo = 0;
    c.root_hi = v_cnt;
    c.counts = h;
    for (int k = 0; k < 8; k++) c.pat_of_k[k] = (k >= 3 && k <= kmax) ? k - 3 : -1;
    c.kmax = kmax;
    c.queue = h + 8;
    c.tier[0] = reinterpret_cast<long long *>(h + kNaHead + kNaMaxV / 2);
    c.tier[1] = c.tier[0] + kNaTierCap;
    c.tier[2] = nullptr;
    c.ovf_n = h + 24;
    c.ovf_cap = kNaTierCap;
    c.bytes = h + 40;
    c.err = reinterpret_cast<unsigned *>(h + 44);
    c.grows = nullptr;  // no HBM tier: vertices with more than 32 * CBIGW neighbours below them are refused
    c.goff = nullptr;
    c.hcap = 32 * gatk::CBIGW;
    c.account = 0;
    c.tl = nullptr;
    c.tl_base = 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(gatk::fill_u64, dim3(1), dim3(64), 0, s, h, kNaHead, 0ull);
    if (v_cnt > 0) {
        hipLaunchKernelGGL(gatk::split_kernel, dim3(1024), dim3(256), 0, s, reinterpret_cast<const long long *>(d_offsets),
                           d_adj, v_cnt, split);
        hipLaunchKernelGGL(gatk::clique_kernel<0>, L.grid, L.block, L.clique_lds, s, c, L.clique_wave_bytes);
        hipLaunchKernelGGL((gatk::clique_block_kernel<gatk::CMIDW, true, 0, gatk::CBLOCKW0, 0>), L.mid_grid,
                           dim3(64 * gatk::CBLOCKW0), L.mid_lds, s, c);
        hipLaunchKernelGGL((gatk::clique_block_kernel<gatk::CBIGW, true, 1, gatk::CBLOCKW1, 0>), L.t1_grid,
                           dim3(64 * gatk::CBLOCKW1), L.t1_lds, s, c);
    }
    HIPCHECK(hipGetLastError());
    unsigned long long raw[8] = {};
    unsigned int err = 0;
    HIPCHECK(hipMemcpyAsync(raw, h, sizeof(raw), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&err, h + 44, sizeof(err), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (err & 16u)
        GAT_FAIL(GAT_ELIMIT, "a vertex has more than %d neighbours below it; the allocation-free count has no HBM tier",
                 32 * gatk::CBIGW);
    if (err & 4u) GAT_FAIL(GAT_ELIMIT, "more than %lld heavy vertices for one workgroup tier", kNaTierCap);
    if (err) GAT_FAIL(GAT_EINTERNAL, "engine reported an internal inconsistency (flags 0x%x)", err);
    for (int k = 3; k <= kmax; k++) counts_out[k - 3] = raw[k - 3];
    return GAT_OK;
}

extern "C" int gat_count_graph(gat_plan *p, const gat_kernel_config *cfg, int32_t replays, uint64_t *counts_out,
                               gat_count_stats *stats) {
    if (!p || replays < 1) GAT_FAIL(GAT_EINVAL, "null plan or replays < 1");
    Launch L;
    int rc = prepare(p, cfg, &L);
    if (rc) return rc;
    PlanDevice *d = p->dev;
    if (d->gexec) {
        (void)hipGraphExecDestroy(d->gexec);
        d->gexec = nullptr;
    }
    if (d->graph) {
        (void)hipGraphDestroy(d->graph);
        d->graph = nullptr;
    }
    HIPCHECK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue(d, L, false, false);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(d->stream, &graph);
    if (rc) return rc;
    if (ec != hipSuccess) GAT_FAIL(GAT_ENODEV, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    d->graph = graph;
    HIPCHECK(hipGraphInstantiate(&d->gexec, graph, nullptr, nullptr, 0));
    for (int r = 0; r < replays; r++) {
        HIPCHECK(hipEventRecord(d->ev[0], d->stream));
        HIPCHECK(hipGraphLaunch(d->gexec, d->stream));
        HIPCHECK(hipEventRecord(d->ev[2], d->stream));
        if (r + 1 < replays) HIPCHECK(hipStreamSynchronize(d->stream));
    }
    rc = finish(p, L, counts_out, stats, false);
    if (rc) return rc;
    if (stats) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, d->ev[0], d->ev[2]));
        stats->seconds = ms / 1000.0;
    }
    return GAT_OK;
}

extern "C" int gat_copy_probe(uint64_t bytes, int32_t reps, double *gbps_out) {
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc) return rc;
    if (bytes < 16 || reps < 1) GAT_FAIL(GAT_EINVAL, "copy probe needs bytes >= 16 and reps >= 1");
    const long long n = (long long)(bytes / 16);
    int4 *src = nullptr, *dst = nullptr;
    HIPCHECK(hipMalloc(&src, n * 16));
    HIPCHECK(hipMalloc(&dst, n * 16));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    const int grid = di.cus * 8;
    hipLaunchKernelGGL(gatk::copy_kernel, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
    HIPCHECK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(gatk::copy_kernel, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
    HIPCHECK(hipEventRecord(e1, nullptr));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    if (gbps_out) *gbps_out = 2.0 * n * 16 * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return GAT_OK;
}
