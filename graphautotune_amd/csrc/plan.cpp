#include "plan.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <sstream>

namespace gat {

using Mask = uint8_t;  // bit l = pattern vertex (level) l
using Perm = std::vector<int>;

// ---------------------------------------------------------------------------
// Unique counting: which restrictions can the engine apply exactly?
//
// The engine counts labeled embeddings psi that satisfy a restriction set R
// (psi(b) < psi(a) for (a,b) in R). For one instance with embedding phi, the
// embeddings are phi o sigma (sigma in Aut), and whether phi o sigma
// satisfies R depends only on the rank order tau of phi's vertex ids. The
// raw count is c * #instances iff every coset tau*Aut holds exactly c rank
// orders that satisfy R.
static uint64_t coset_factor(const Pattern &q, const std::vector<Perm> &G, const std::vector<Pair> &R) {
    const int n = q.n;
    Perm tau(n);
    std::iota(tau.begin(), tau.end(), 0);
    int64_t nfact = 1;
    for (int i = 2; i <= n; i++) nfact *= i;
    const int64_t cosets = nfact / (int64_t)G.size();
    std::map<uint64_t, int64_t> per_coset;
    int64_t work = 0;
    do {
        bool ok = true;
        for (auto [a, b] : R)
            if (!(tau[b] < tau[a])) {
                ok = false;
                break;
            }
        if (!ok) continue;
        if ((work += (int64_t)G.size()) > 50000000) return 0;  // too costly: treat as non-constant
        uint64_t rep = ~0ull;
        for (const auto &s : G) {
            uint64_t e = 0;
            for (int i = 0; i < n; i++) e = e * 16 + (uint64_t)tau[s[i]];
            rep = std::min(rep, e);
        }
        per_coset[rep]++;
    } while (std::next_permutation(tau.begin(), tau.end()));
    if ((int64_t)per_coset.size() != cosets) return 0;
    const int64_t c = per_coset.begin()->second;
    for (auto &kv : per_coset)
        if (kv.second != c) return 0;
    return (uint64_t)c;
}

uint64_t unique_restrictions(const Schedule &s, std::vector<Pair> *used) {
    const Pattern &q = s.ordered;
    const std::vector<Perm> G = automorphisms(q);
    const int b = s.basic;
    std::vector<Pair> no_ss, basic_only;
    for (auto pr : s.chosen) {
        if (pr.first < b) no_ss.push_back(pr);  // first < second, so first >= b means both in the suffix
        if (pr.second < b) basic_only.push_back(pr);
    }
    std::vector<std::vector<Pair>> cands{no_ss};
    if (basic_only != no_ss) cands.push_back(basic_only);
    for (auto &R : cands) {
        if (R.empty()) continue;
        const uint64_t c = coset_factor(q, G, R);
        if (c) {
            *used = R;
            return c;
        }
    }
    used->clear();
    return (uint64_t)G.size();
}

// ---------------------------------------------------------------------------
namespace {

struct HSlot {
    int kind = GATP_LIST;
    int node = 0, depth = 0;
    std::vector<int> data;
    Mask T = 0, X = 0;  // COUNT: exact bound levels and exclusions
    int father = -1;
    // truncation uses: (type, group or slot)
    std::vector<std::pair<int, int>> uses;  // type 0 = group, 1 = father-of slot, 2 = none
    bool resolved = false, notrunc = false;
    std::vector<Mask> groups;
};

struct HNode {
    int depth = 0, parent = -1;
    Mask L = 0, B = 0, X = 0, A = 0;
    int loop_slot = -1;
    std::vector<int> children, ops, terms;
};

struct HVid {
    int slot;
    int mode;
    Mask bound = 0, excl = 0;
    int refn = 0;
    bool operator==(const HVid &o) const {
        return slot == o.slot && mode == o.mode && bound == o.bound && excl == o.excl && refn == o.refn;
    }
};

struct HTerm {
    int pattern, node;
    std::vector<HVid> vids;
    std::vector<GatSub> subs;  // vid = index into this term's vids
};

struct Compiler {
    std::vector<HNode> nodes;
    std::vector<HSlot> slots;
    std::vector<HTerm> terms;

    Compiler() { nodes.push_back(HNode{}); }

    int child(int parent, int depth, Mask L, Mask B, Mask X, Mask A) {
        for (int c : nodes[parent].children) {
            const HNode &h = nodes[c];
            if (h.L == L && h.B == B && h.X == X && h.A == A) return c;
        }
        HNode h;
        h.depth = depth;
        h.parent = parent;
        h.L = L;
        h.B = B;
        h.X = X;
        h.A = A;
        nodes.push_back(h);
        nodes[parent].children.push_back((int)nodes.size() - 1);
        return (int)nodes.size() - 1;
    }

    // ALIAS or LIST slot holding the intersection of N(v_i), i in data.
    int list_slot(const std::vector<int> &path, const std::vector<int> &data) {
        const int node = path[data.back()];
        for (size_t i = 0; i < slots.size(); i++)
            if (slots[i].kind != GATP_COUNT && slots[i].node == node && slots[i].data == data) return (int)i;
        HSlot s;
        s.kind = data.size() == 1 ? GATP_ALIAS : GATP_LIST;
        s.node = node;
        s.depth = data.back();
        s.data = data;
        if (data.size() > 1) {
            std::vector<int> fd(data.begin(), data.end() - 1);
            s.father = list_slot(path, fd);
        }
        slots.push_back(s);
        const int id = (int)slots.size() - 1;
        if (slots[id].father >= 0) slots[slots[id].father].uses.push_back({1, id});
        nodes[node].ops.push_back(id);
        return id;
    }

    int count_slot(const std::vector<int> &path, const std::vector<int> &data, Mask T, Mask X) {
        const int node = path[data.back()];
        for (size_t i = 0; i < slots.size(); i++)
            if (slots[i].kind == GATP_COUNT && slots[i].node == node && slots[i].data == data &&
                slots[i].T == T && slots[i].X == X)
                return (int)i;
        HSlot s;
        s.kind = GATP_COUNT;
        s.node = node;
        s.depth = data.back();
        s.data = data;
        s.T = T;
        s.X = X;
        if (data.size() > 1) {
            std::vector<int> fd(data.begin(), data.end() - 1);
            s.father = list_slot(path, fd);
        }
        s.resolved = true;
        if (T)
            s.groups.push_back(T);
        else
            s.notrunc = true;
        slots.push_back(s);
        const int id = (int)slots.size() - 1;
        if (slots[id].father >= 0) slots[slots[id].father].uses.push_back({1, id});
        nodes[node].ops.push_back(id);
        return id;
    }

    void resolve(int id) {
        HSlot &s = slots[id];
        if (s.resolved) return;
        const Mask lim = Mask((1u << (s.depth + 1)) - 1);
        std::vector<Mask> g;
        bool none = false;
        for (auto [type, val] : s.uses) {
            if (type == 2) {
                none = true;
            } else if (type == 0) {
                const Mask m = Mask(val) & lim;
                if (!m) none = true; else g.push_back(m);
            } else {
                resolve(val);
                const HSlot &c = slots[val];
                if (c.notrunc) none = true;
                for (Mask cg : c.groups) {
                    const Mask m = cg & lim;
                    if (!m) none = true; else g.push_back(m);
                }
            }
        }
        s.resolved = true;
        if (none || g.empty()) {
            s.notrunc = true;
            return;
        }
        // keep minimal groups: a superset group gives a smaller bound and
        // never decides the maximum
        std::sort(g.begin(), g.end());
        g.erase(std::unique(g.begin(), g.end()), g.end());
        std::vector<Mask> keep;
        for (Mask x : g) {
            bool dominated = false;
            for (Mask y : g)
                if (y != x && (y & x) == y) dominated = true;
            if (!dominated) keep.push_back(x);
        }
        while (keep.size() > 4) {  // merge (weaker bound) until it fits
            Mask m = keep[keep.size() - 1] & keep[keep.size() - 2];
            keep.pop_back();
            keep.pop_back();
            if (!m) {
                s.notrunc = true;
                return;
            }
            keep.push_back(m);
        }
        s.groups = keep;
    }
};

}  // namespace

int compile_plan(const std::vector<std::string> &adj, int mode, const PlanOptions &opt, CompiledPlan *out) {
    if (adj.empty()) GAT_FAIL(GAT_EINVAL, "no patterns");
    if (mode != GAT_MODE_UNIQUE && mode != GAT_MODE_REFERENCE && mode != GAT_MODE_LABELED)
        GAT_FAIL(GAT_EINVAL, "unknown count mode %d", mode);
    if (opt.schedule_policy != GAT_SCHEDULE_REFERENCE && opt.schedule_policy != GAT_SCHEDULE_GRAPH &&
        opt.schedule_policy != GAT_SCHEDULE_RANKED)
        GAT_FAIL(GAT_EINVAL, "unknown schedule policy %d", opt.schedule_policy);
    if (opt.schedule_policy != GAT_SCHEDULE_REFERENCE && mode == GAT_MODE_REFERENCE)
        GAT_FAIL(GAT_EINVAL, "reference-mode counts are defined only for the reference's schedule");
    if (opt.schedule_rank < 0) GAT_FAIL(GAT_EINVAL, "negative schedule rank");
    if (!opt.schedule_ranks.empty() && opt.schedule_ranks.size() != adj.size())
        GAT_FAIL(GAT_EINVAL, "%zu schedule ranks for %zu patterns", opt.schedule_ranks.size(), adj.size());
    for (int r : opt.schedule_ranks)
        if (r < 0) GAT_FAIL(GAT_EINVAL, "negative schedule rank");
    if ((int)adj.size() > GATP_MAX_TERMS) GAT_FAIL(GAT_ELIMIT, "at most %d patterns per plan", GATP_MAX_TERMS);
    CompiledPlan cp;
    cp.relabel = mode != GAT_MODE_REFERENCE;
    Compiler C;
    for (size_t pi = 0; pi < adj.size(); pi++) {
        PatternInfo info;
        info.adj = adj[pi];
        info.mode = mode;
        Pattern p;
        int rc = parse_pattern(adj[pi], &p);
        if (rc) return rc;
        if (opt.schedule_policy == GAT_SCHEDULE_REFERENCE) {
            rc = build_schedule(p, &info.sched);
        } else {
            std::vector<ScheduleCandidate> cands;
            rc = schedule_candidates(p, opt.stats, &cands);
            if (!rc)
            {
                const int rank = opt.schedule_ranks.empty() ? opt.schedule_rank : opt.schedule_ranks[pi];
                rc = build_schedule_from(p, cands[std::min<size_t>((size_t)rank, cands.size() - 1)], &info.sched);
            }
        }
        if (rc) return rc;
        const Schedule &s = info.sched;
        const Pattern &q = s.ordered;
        const int b = s.basic, n = s.n;
        if (b > GATP_MAX_DEPTH) GAT_FAIL(GAT_ELIMIT, "pattern %zu has %d basic vertices (> %d)", pi, b, GATP_MAX_DEPTH);
        const bool ref = mode == GAT_MODE_REFERENCE;
        if (ref)
            for (int l = 1; l < b; l++) {
                const int pid = s.loop_set_prefix[l];
                // The reference reads such a loop set from a buffer that was
                // only given a size (engine_kernel.cuh:153-154 after
                // :276-282): its result is undefined.
                if (s.size_only[pid] && s.prefixes[pid].data.size() > 1)
                    GAT_FAIL(GAT_EUNSUPPORTED,
                             "reference engine behaviour is undefined for pattern %s: the loop set of "
                             "vertex %d is a size-only prefix",
                             adj[pi].c_str(), l);
            }

        std::vector<Pair> R;
        if (ref) {
            R = s.filtered;
            info.divisor = 1;
        } else {
            info.divisor = unique_restrictions(s, &R);
            if (mode == GAT_MODE_LABELED) info.multiplier = (uint64_t)s.automorphism_count;
        }
        info.used_restrictions = R;
        Mask bound_of[kMaxV] = {};
        for (auto [f, t] : R) bound_of[t] |= Mask(1u << f);

        // ---- trie path ----
        std::vector<int> path(b, 0);
        Mask Lm[kMaxV] = {};
        for (int l = 1; l < b; l++) {
            Mask L = Mask(q.row[l] & ((1u << l) - 1));
            Lm[l] = L;
            const Mask earlier = Mask((1u << l) - 1);
            Mask B = bound_of[l];
            Mask X = 0, A = 0;
            if (ref)
                A = Mask(earlier & ~L & ~B);  // reference: duplicates descend with emptied sets
            else
                X = Mask(earlier & ~L & ~B);  // unique: duplicates are skipped
            path[l] = C.child(path[l - 1], l, L, B, X, A);
        }
        auto data_of = [](Mask m) {
            std::vector<int> d;
            for (int i = 0; i < 8; i++)
                if (m >> i & 1) d.push_back(i);
            return d;
        };
        // loop sets
        for (int l = 1; l < b; l++) {
            const int ls = C.list_slot(path, data_of(Lm[l]));
            HNode &h = C.nodes[path[l]];
            if (h.loop_slot >= 0 && h.loop_slot != ls) GAT_FAIL(GAT_EINVAL, "internal: loop slot conflict");
            h.loop_slot = ls;
            C.slots[ls].uses.push_back({0, h.B});
        }
        // ---- IEP terminal at the last basic level ----
        HTerm term;
        term.pattern = (int)pi;
        term.node = path[b - 1];
        auto add_vid = [&](const HVid &v) {
            for (size_t i = 0; i < term.vids.size(); i++)
                if (term.vids[i] == v) return (int)i;
            term.vids.push_back(v);
            return (int)term.vids.size() - 1;
        };
        if (ref) {
            // Reference semantics (engine_kernel.cuh:343-432): size-only
            // prefixes give |set| computed at their own level excluding the
            // vertices chosen before it (array_vertex_set.cuh:248-275);
            // other prefixes give |set| minus the chosen vertices found in
            // it (array_vertex_set.cuh:318-331); a depth-1 prefix is the
            // neighbour list itself (engine_kernel.cuh:264-266).
            std::vector<int> vid_map;
            for (int pid : s.iep_vertex_id) {
                const std::vector<int> &d = s.prefixes[pid].data;
                HVid v;
                if (d.size() == 1) {
                    v.slot = C.list_slot(path, d);
                    if (s.size_only[pid]) {
                        v.mode = GATP_VID_LIST_BOUNDED;  // plain length
                        C.slots[v.slot].uses.push_back({2, 0});
                    } else {
                        v.mode = GATP_VID_LIST_REF;
                        v.refn = b;
                        C.slots[v.slot].uses.push_back({2, 0});
                    }
                } else if (s.size_only[pid]) {
                    Mask Xq = Mask(((1u << d.back()) - 1));
                    for (int i : d) Xq &= Mask(~(1u << i));
                    v.slot = C.count_slot(path, d, 0, Xq);
                    v.mode = GATP_VID_COUNT;
                } else {
                    v.slot = C.list_slot(path, d);
                    v.mode = GATP_VID_LIST_REF;
                    v.refn = b;
                    C.slots[v.slot].uses.push_back({2, 0});
                }
                vid_map.push_back(add_vid(v));
            }
            for (size_t g = 0; g < s.iep_ans_pos.size(); g++)
                term.subs.push_back(GatSub{int16_t(vid_map[s.iep_ans_pos[g]]), int16_t(s.iep_flag[g]), s.iep_coef[g]});
        } else {
            const int base = n - s.iep_suffix;
            for (const auto &grp : iep_groups(s.iep_suffix)) {
                for (size_t k = 0; k < grp.blocks.size(); k++) {
                    Mask D = 0, T = 0;
                    for (int sv : grp.blocks[k]) {
                        D |= Mask(q.row[base + sv] & ((1u << base) - 1));
                        T |= bound_of[base + sv];
                    }
                    const std::vector<int> d = data_of(D);
                    const Mask all = Mask((1u << b) - 1);
                    const Mask X = Mask(all & ~D & ~T);
                    HVid v;
                    if (d.back() == b - 1) {
                        v.slot = C.count_slot(path, d, T, X);
                        v.mode = GATP_VID_COUNT;
                    } else {
                        v.slot = C.list_slot(path, d);
                        v.mode = GATP_VID_LIST_BOUNDED;
                        v.bound = T;
                        v.excl = X;
                        C.slots[v.slot].uses.push_back(T ? std::make_pair(0, (int)T) : std::make_pair(2, 0));
                    }
                    const bool last = k + 1 == grp.blocks.size();
                    term.subs.push_back(GatSub{int16_t(add_vid(v)), int16_t(last ? 1 : 0), last ? grp.coef : 0});
                }
            }
        }
        C.nodes[term.node].terms.push_back((int)C.terms.size());
        C.terms.push_back(std::move(term));
        cp.patterns.push_back(std::move(info));
    }
    // A COUNT whose prefix is also materialised as a LIST at the same node is
    // read off that list (lower bound + exclusions) instead of being
    // intersected a second time; batched plans hit this when one pattern
    // counts the set another pattern loops over (K4 + K5).
    std::vector<int> dead(C.slots.size(), 0);
    for (size_t i = 0; i < C.slots.size(); i++) {
        HSlot &cs = C.slots[i];
        if (cs.kind != GATP_COUNT) continue;
        int list = -1;
        for (size_t j = 0; j < C.slots.size(); j++)
            if (C.slots[j].kind == GATP_LIST && C.slots[j].node == cs.node && C.slots[j].data == cs.data) list = (int)j;
        if (list < 0) continue;
        for (HTerm &t : C.terms)
            for (HVid &v : t.vids)
                if (v.slot == (int)i && v.mode == GATP_VID_COUNT) {
                    v.slot = list;
                    v.mode = GATP_VID_LIST_BOUNDED;
                    v.bound = cs.T;
                    v.excl = cs.X;
                    C.slots[list].uses.push_back(cs.T ? std::make_pair(0, (int)cs.T) : std::make_pair(2, 0));
                }
        dead[i] = 1;
        if (cs.father >= 0) {  // drop the father-use it registered
            auto &u = C.slots[cs.father].uses;
            u.erase(std::remove(u.begin(), u.end(), std::make_pair(1, (int)i)), u.end());
        }
        auto &ops = C.nodes[cs.node].ops;
        ops.erase(std::remove(ops.begin(), ops.end(), (int)i), ops.end());
    }
    for (size_t i = 0; i < C.slots.size(); i++)
        if (!dead[i]) C.resolve((int)i);

    // ---- emit ----
    if ((int)C.nodes.size() > GATP_MAX_NODES) GAT_FAIL(GAT_ELIMIT, "plan needs %zu nodes", C.nodes.size());
    if ((int)C.slots.size() > GATP_MAX_SLOTS) GAT_FAIL(GAT_ELIMIT, "plan needs %zu slots", C.slots.size());
    GatPlanDev &d = cp.dev;
    std::memset(&d, 0, sizeof(d));
    // BFS order so that children are consecutive
    std::vector<int> order{0}, newid(C.nodes.size(), -1);
    for (size_t i = 0; i < order.size(); i++)
        for (int c : C.nodes[order[i]].children) order.push_back(c);
    for (size_t i = 0; i < order.size(); i++) newid[order[i]] = (int)i;
    // storage regions per depth for LIST slots
    int maxd = 0;
    for (auto &h : C.nodes) maxd = std::max(maxd, h.depth);
    std::vector<int> per_depth(maxd + 1, 0);
    // A neighbour list that later ops search (it is some op's father) gets an
    // LDS buffer too: it is staged there when it fits.
    std::vector<char> searched(C.slots.size(), 0);
    for (size_t i = 0; i < C.slots.size(); i++)
        if (!dead[i] && C.slots[i].father >= 0) searched[C.slots[i].father] = 1;
    auto needs_store = [&](int sid) {
        return C.slots[sid].kind == GATP_LIST || (C.slots[sid].kind == GATP_ALIAS && searched[sid]);
    };
    for (auto &h : C.nodes) {
        int k = 0;
        for (int sid : h.ops)
            if (needs_store(sid)) k++;
        per_depth[h.depth] = std::max(per_depth[h.depth], k);
    }
    std::vector<int> store_base(maxd + 2, 0);
    for (int l = 0; l <= maxd; l++) store_base[l + 1] = store_base[l] + per_depth[l];
    d.n_store = store_base[maxd + 1];
    if (d.n_store > GATP_MAX_STORE) GAT_FAIL(GAT_ELIMIT, "plan needs %d set buffers (> %d)", d.n_store, GATP_MAX_STORE);

    d.n_slots = (int)C.slots.size();
    int nops = 0, nterms = 0, nvids = 0, nsubs = 0, max_leaf = 0;
    for (size_t ii = 0; ii < order.size(); ii++) {
        const HNode &h = C.nodes[order[ii]];
        GatNode &g = d.nodes[ii];
        g.depth = int8_t(h.depth);
        g.bound = h.B;
        g.excl = h.X;
        g.ap_levels = h.A;
        g.parent = int16_t(h.parent < 0 ? -1 : newid[h.parent]);
        g.loop_slot = int16_t(h.loop_slot);
        g.leaf = h.children.empty();
        g.child_count = int16_t(h.children.size());
        g.child_begin = int16_t(h.children.empty() ? -1 : newid[h.children[0]]);
        g.next_sibling = -1;
        if (h.parent >= 0) {
            const auto &sib = C.nodes[h.parent].children;
            auto it = std::find(sib.begin(), sib.end(), order[ii]);
            if (it + 1 != sib.end()) g.next_sibling = int16_t(newid[*(it + 1)]);
        }
        g.op_begin = int16_t(nops);
        int lists = 0, leafc = 0;
        for (int sid : h.ops) {
            if (nops >= GATP_MAX_OPS) GAT_FAIL(GAT_ELIMIT, "plan needs more than %d ops", GATP_MAX_OPS);
            HSlot &s = C.slots[sid];
            GatOp &o = d.ops[nops++];
            o.slot = int16_t(sid);
            o.father = int16_t(s.father);
            o.kind = int8_t(s.kind);
            o.ngroups = int8_t(s.notrunc ? 0 : s.groups.size());
            for (size_t k = 0; k < s.groups.size() && !s.notrunc; k++) o.groups[k] = s.groups[k];
            o.excl = s.X;
            o.leaf_idx = -1;
            o.store = -1;
            GatSlot &gs = d.slots[sid];
            gs.kind = int8_t(s.kind);
            gs.depth = int8_t(s.depth);
            gs.node = int16_t(ii);
            gs.leaf_idx = -1;
            gs.store = -1;
            if (needs_store(sid)) o.store = gs.store = int8_t(store_base[h.depth] + lists++);
            if ((s.kind == GATP_COUNT || s.kind == GATP_ALIAS) && g.leaf) o.leaf_idx = gs.leaf_idx = int8_t(leafc++);
        }
        if (leafc > GATP_MAX_LEAF_OPS) GAT_FAIL(GAT_ELIMIT, "leaf needs %d counters (> %d)", leafc, GATP_MAX_LEAF_OPS);
        max_leaf = std::max(max_leaf, leafc);
        g.op_count = int16_t(nops - g.op_begin);
        g.term_begin = int16_t(nterms);
        for (int t : h.terms) {
            const HTerm &ht = C.terms[t];
            if (nvids + (int)ht.vids.size() > GATP_MAX_VIDS || nsubs + (int)ht.subs.size() > GATP_MAX_SUBS)
                GAT_FAIL(GAT_ELIMIT, "plan IEP tables too large");
            GatTerm &gt = d.terms[nterms++];
            gt.pattern = int16_t(ht.pattern);
            gt.vid_begin = int16_t(nvids);
            gt.vid_count = int16_t(ht.vids.size());
            for (const HVid &v : ht.vids) {
                GatVid &gv = d.vids[nvids++];
                gv.slot = int16_t(v.slot);
                gv.mode = int8_t(v.mode);
                gv.bound = v.bound;
                gv.excl = v.excl;
                gv.refn = uint8_t(v.refn);
            }
            gt.sub_begin = int16_t(nsubs);
            gt.sub_count = int16_t(ht.subs.size());
            for (const GatSub &sb : ht.subs) {
                GatSub &o = d.subs[nsubs++];
                o = sb;
                o.vid = int16_t(gt.vid_begin + sb.vid);
            }
        }
        g.term_count = int16_t(nterms - g.term_begin);
    }
    d.n_nodes = (int)order.size();
    d.n_ops = nops;
    d.n_terms = nterms;
    d.n_vids = nvids;
    d.n_subs = nsubs;
    d.n_patterns = (int)adj.size();
    d.max_leaf_ops = max_leaf;
    d.max_depth = maxd;
    // bitmap eligibility (clique-only plans)
    if (mode != GAT_MODE_REFERENCE) {
        uint32_t ks = 0;
        bool all = true;
        for (size_t i = 0; i < cp.patterns.size(); i++) {
            const Pattern &p = cp.patterns[i].sched.input;
            const bool clique = p.edge_count() == p.n * (p.n - 1) / 2;
            if (!clique || p.n > 6) {
                all = false;
                break;
            }
            if (cp.clique_pattern[p.n] >= 0) {  // the same clique twice: leave it to the array store
                all = false;
                break;
            }
            cp.clique_pattern[p.n] = (int)i;
            ks |= 1u << p.n;
        }
        if (all) cp.clique_ks = ks;
    }
    d.root_slot = -1;
    for (int sid : C.nodes[0].ops)
        if (C.slots[sid].data == std::vector<int>{0}) d.root_slot = sid;
    if (d.root_slot < 0) GAT_FAIL(GAT_EINVAL, "internal: no root slot");
    *out = std::move(cp);
    return GAT_OK;
}

std::string describe_plan(const GatPlanDev &d) {
    std::ostringstream os;
    static const char *kinds[] = {"ALIAS", "LIST", "COUNT"};
    os << "nodes " << d.n_nodes << " slots " << d.n_slots << " ops " << d.n_ops << " terms " << d.n_terms
       << " store " << d.n_store << " max_leaf_ops " << d.max_leaf_ops << "\n";
    for (int i = 0; i < d.n_nodes; i++) {
        const GatNode &g = d.nodes[i];
        os << "node " << i << " depth " << int(g.depth) << " parent " << g.parent << " loop " << g.loop_slot
           << " bound 0x" << std::hex << int(g.bound) << " excl 0x" << int(g.excl) << " ap 0x" << int(g.ap_levels)
           << std::dec << (g.leaf ? " leaf" : "") << " children " << g.child_begin << "+" << g.child_count
           << " sib " << g.next_sibling << "\n";
        for (int k = 0; k < g.op_count; k++) {
            const GatOp &o = d.ops[g.op_begin + k];
            os << "   op slot " << o.slot << " " << kinds[o.kind] << " father " << o.father << " groups[";
            for (int j = 0; j < o.ngroups; j++) os << std::hex << "0x" << int(o.groups[j]) << std::dec << " ";
            os << "] excl 0x" << std::hex << int(o.excl) << std::dec << " store " << int(o.store) << " leaf_idx "
               << int(o.leaf_idx) << "\n";
        }
        for (int t = 0; t < g.term_count; t++) {
            const GatTerm &tm = d.terms[g.term_begin + t];
            os << "   term pattern " << tm.pattern << " vids:";
            for (int v = 0; v < tm.vid_count; v++) {
                const GatVid &gv = d.vids[tm.vid_begin + v];
                os << " (slot " << gv.slot << " mode " << int(gv.mode) << " bound 0x" << std::hex << int(gv.bound)
                   << " excl 0x" << int(gv.excl) << std::dec << " refn " << int(gv.refn) << ")";
            }
            os << " subs:";
            for (int k = 0; k < tm.sub_count; k++) {
                const GatSub &sb = d.subs[tm.sub_begin + k];
                os << " " << sb.vid << (sb.flag ? "!" : "") << "*" << sb.coef;
            }
            os << "\n";
        }
    }
    return os.str();
}

}  // namespace gat
