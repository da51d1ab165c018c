// Pattern graphs and the Schedule builder (restates include/core/pattern.hpp
// and include/core/schedule.hpp of the reference; decisions are identical,
// data structures are bitmask based).
#include "pattern.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>

namespace gat {

int Pattern::edge_count() const {
    int e = 0;
    for (int i = 0; i < n; i++) e += __builtin_popcount(row[i]);
    return e / 2;
}

int Pattern::max_degree() const {
    int m = 0;
    for (int i = 0; i < n; i++) m = std::max(m, __builtin_popcount(row[i]));
    return m;
}

Pattern Pattern::relabel(const int *perm) const {
    Pattern q;
    q.n = n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            if (has_edge(i, j)) q.add_edge(perm[i], perm[j]);
    return q;
}

bool Pattern::operator==(const Pattern &o) const {
    if (n != o.n) return false;
    for (int i = 0; i < n; i++)
        if (row[i] != o.row[i]) return false;
    return true;
}

std::string Pattern::str() const {
    std::string s(n * n, '0');
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            if (has_edge(i, j)) s[i * n + j] = '1';
    return s;
}

int parse_pattern(const std::string &s, Pattern *out) {
    const int len = (int)s.size();
    int n = 0;
    while ((n + 1) * (n + 1) <= len) n++;
    if (n * n != len || n < 1)
        GAT_FAIL(GAT_EINVAL, "pattern string length %d is not a square", len);
    if (n > kMaxV) GAT_FAIL(GAT_ELIMIT, "pattern has %d vertices; at most %d supported", n, kMaxV);
    for (char c : s)
        if (c != '0' && c != '1')
            GAT_FAIL(GAT_EINVAL, "pattern string may only hold '0' and '1'");
    Pattern p;
    p.n = n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++)
            if (s[i * n + j] == '1' || s[j * n + i] == '1') p.add_edge(i, j);
    *out = p;
    return GAT_OK;
}

bool order_connected(const Pattern &p) {
    for (int i = 1; i < p.n; i++)
        if ((p.row[i] & ((1u << i) - 1)) == 0) return false;
    return true;
}

int iep_suffix_length(const Pattern &p) {
    for (int k = 2; k <= p.n - 2; k++) {
        const int v = p.n - k;
        // any edge from v to a later vertex ends the independent suffix
        if (p.row[v] >> (v + 1)) return k - 1;
    }
    return p.n - 2;
}

std::vector<std::vector<int>> automorphisms(const Pattern &p) {
    std::vector<std::vector<int>> out;
    std::vector<int> perm(p.n);
    std::iota(perm.begin(), perm.end(), 0);
    do {
        if (p.relabel(perm.data()) == p) out.push_back(perm);
    } while (std::next_permutation(perm.begin(), perm.end()));
    return out;
}

// Cycle decomposition, each cycle started at its smallest unused element
// (schedule.hpp:193-213).
static std::vector<std::vector<int>> cycles_of(const std::vector<int> &perm) {
    const int n = (int)perm.size();
    std::vector<char> used(n, 0);
    std::vector<std::vector<int>> res;
    for (int i = 0; i < n; i++) {
        if (used[i]) continue;
        std::vector<int> c{i};
        used[i] = 1;
        for (int x = perm[i]; !used[x]; x = perm[x]) {
            used[x] = 1;
            c.push_back(x);
        }
        res.push_back(std::move(c));
    }
    return res;
}

int64_t linear_extensions(int n, const std::vector<Pair> &pairs) {
    // pred[b] = set of a that must come before b
    uint32_t pred[32] = {};
    for (auto [a, b] : pairs) pred[b] |= 1u << a;
    std::vector<int64_t> f(size_t(1) << n, 0);
    f[0] = 1;
    for (uint32_t m = 0; m < (1u << n); m++) {
        if (!f[m]) continue;
        for (int x = 0; x < n; x++)
            if (!(m >> x & 1) && (pred[x] & ~m) == 0) f[m | (1u << x)] += f[m];
    }
    return f[(size_t(1) << n) - 1];
}

static bool acyclic(const std::vector<Pair> &pairs, int n) {
    uint32_t pred[32] = {};
    for (auto [a, b] : pairs) pred[b] |= 1u << a;
    uint32_t done = 0;
    for (bool progress = true; progress;) {
        progress = false;
        for (int x = 0; x < n; x++)
            if (!(done >> x & 1) && (pred[x] & ~done) == 0) {
                done |= 1u << x;
                progress = true;
            }
    }
    return done == (n == 32 ? ~0u : (1u << n) - 1);
}

using Perm = std::vector<int>;
using Cycles = std::vector<std::vector<int>>;

// schedule.hpp:239-293. The reference never increments its early-exit
// counter inside this function, so every 2-cycle of every remaining
// automorphism is branched on; the recursion order below is the same.
// A subtree's output depends only on (set of base pairs, remaining
// automorphisms), and a restriction set found again later can never win
// (ties keep the first candidate, and the estimate sorts the pairs), so a
// state seen before is skipped: the distinct sets still come out in the
// reference's first-occurrence order.
using DfsKey = std::pair<std::vector<Pair>, std::vector<Perm>>;
static void restriction_dfs(const std::vector<Pair> &base, std::vector<Perm> iso,
                            std::vector<Cycles> groups, std::vector<std::vector<Pair>> &out,
                            int n, std::set<DfsKey> &seen) {
    {
        DfsKey key{base, iso};
        std::sort(key.first.begin(), key.first.end());
        key.first.erase(std::unique(key.first.begin(), key.first.end()), key.first.end());
        std::sort(key.second.begin(), key.second.end());
        if (!seen.insert(std::move(key)).second) return;
    }
    std::vector<Perm> kept_iso;
    std::vector<Cycles> kept_groups;
    for (size_t i = 0; i < iso.size(); i++) {
        std::vector<Pair> test = base;
        for (auto [a, b] : base) test.emplace_back(iso[i][a], iso[i][b]);
        if (acyclic(test, n)) {
            kept_groups.push_back(groups[i]);
            kept_iso.push_back(iso[i]);
        }
    }
    if (kept_iso.size() == 1) {
        out.push_back(base);
        return;
    }
    for (size_t i = 0; i < kept_groups.size(); i++) {
        for (const auto &c : kept_groups[i]) {
            if (c.size() != 2) continue;
            std::vector<Cycles> ng = kept_groups;
            std::vector<Perm> ni = kept_iso;
            ng.erase(ng.begin() + i);
            ni.erase(ni.begin() + i);
            std::vector<Pair> nb = base;
            nb.emplace_back(c[0], c[1]);
            restriction_dfs(nb, ni, ng, out, n, seen);
        }
    }
}

static std::vector<std::vector<Pair>> restriction_candidates_from(const Pattern &p,
                                                                  const std::vector<Perm> &aut) {
    // schedule.hpp:296-338: automorphisms that are a single transposition
    // contribute their pair directly; the others go to the search.
    std::vector<Cycles> groups;
    for (const auto &a : aut) groups.push_back(cycles_of(a));
    std::vector<Perm> rest_iso;
    std::vector<Cycles> rest_groups;
    std::vector<Pair> pairs;
    for (size_t i = 0; i < groups.size(); i++) {
        int twos = 0;
        Pair found{-1, -1};
        for (const auto &c : groups[i]) {
            if (c.size() == 2) {
                twos++;
                found = {c[0], c[1]};
            } else if (c.size() != 1) {
                twos = -1;
                break;
            }
        }
        if (twos == 1) {
            pairs.push_back(found);
        } else {
            rest_groups.push_back(groups[i]);
            rest_iso.push_back(aut[i]);
        }
    }
    std::vector<std::vector<Pair>> candidates;
    std::set<DfsKey> seen;
    restriction_dfs(pairs, rest_iso, rest_groups, candidates, p.n, seen);
    // schedule.hpp:340-360: keep the sets whose linear-extension count in
    // the complete graph equals n!/|Aut|.
    const int64_t want = linear_extensions(p.n, {}) / (int64_t)aut.size();
    std::vector<std::vector<Pair>> valid;
    for (auto &r : candidates)
        if (linear_extensions(p.n, r) == want) valid.push_back(r);
    return valid;
}

std::vector<std::vector<Pair>> restriction_candidates(const Pattern &p) {
    return restriction_candidates_from(p, automorphisms(p));
}

double estimate_restrictions(const Pattern &p, const std::vector<Pair> &restrictions, int v_cnt,
                             int64_t e_cnt, int64_t tri_cnt) {
    // Arithmetic is kept in the reference's evaluation order
    // (schedule.hpp:362-437) so the doubles, and therefore the tie-breaks
    // between candidates, are bit-identical.
    const int size = p.n;
    const int iep_suffix_num = iep_suffix_length(p);
    std::vector<Pair> r = restrictions;
    const double p0 = e_cnt * 1.0 / v_cnt / v_cnt;
    const double p1 = tri_cnt * 1.0 * v_cnt / e_cnt / e_cnt;
    const int rs = (int)r.size();
    std::sort(r.begin(), r.end());
    // sum[i]: permutations tmp with tmp[first] > tmp[second] for the first
    // i+1 sorted restrictions, i.e. linear extensions of the reversed pairs.
    std::vector<double> sum(rs);
    for (int i = 0; i < rs; i++) {
        std::vector<Pair> rev;
        for (int k = 0; k <= i; k++) rev.emplace_back(r[k].second, r[k].first);
        sum[i] = (double)linear_extensions(size, rev);
    }
    double total = 1;
    for (int i = 2; i <= size; i++) total *= i;
    for (int i = 0; i < rs; i++) sum[i] = sum[i] / total;
    for (int i = rs - 1; i > 0; --i) sum[i] /= sum[i - 1];

    std::vector<std::vector<int>> invariant_size(size);
    double val = 1;
    for (int i = size - iep_suffix_num - 1; i >= 0; --i) {
        int cnt_forward = 0;
        for (int j = 0; j < i; j++)
            if (p.has_edge(j, i)) ++cnt_forward;
        int c = cnt_forward;
        for (int j = i - 1; j >= 0; --j)
            if (p.has_edge(j, i)) invariant_size[j].push_back(c--);
        for (size_t j = 0; j < invariant_size[i].size(); j++)
            if (invariant_size[i][j] > 1)
                val += v_cnt * p0 * std::pow(p1, (double)(invariant_size[i][j] - 2)) *
                       std::log2(p0 * v_cnt);
        for (int j = 0; j < rs; j++)
            if (r[j].second == i) val *= sum[j];
        if (i)
            val *= v_cnt * p0 * std::pow(p1, (double)(cnt_forward - 1));
        else
            val *= v_cnt;
    }
    return val;
}

int iep_block_coefficient(int k) {
    int f = 1;
    for (int i = 2; i < k; i++) f *= i;
    return (k % 2 == 1) ? f : -f;
}

static void iep_groups_rec(int depth, std::vector<int> &id, int group_cnt, int suffix,
                           std::vector<IepGroup> &out) {
    if (depth == suffix) {
        IepGroup g;
        g.coef = 1;
        for (int gid = 0; gid < group_cnt; gid++) {
            std::vector<int> cur;
            for (int v = 0; v < suffix; v++)
                if (id[v] == gid) cur.push_back(v);
            g.coef *= iep_block_coefficient((int)cur.size());
            g.blocks.push_back(std::move(cur));
        }
        out.push_back(std::move(g));
        return;
    }
    id[depth] = group_cnt;  // open a new block first
    iep_groups_rec(depth + 1, id, group_cnt + 1, suffix, out);
    for (int i = 0; i < group_cnt; i++) {  // then join an existing one
        id[depth] = i;
        iep_groups_rec(depth + 1, id, group_cnt, suffix, out);
    }
}

std::vector<IepGroup> iep_groups(int suffix) {
    std::vector<IepGroup> out;
    std::vector<int> id(std::max(suffix, 0));
    iep_groups_rec(0, id, 0, suffix, out);
    return out;
}

namespace {

struct Builder {
    Schedule &s;
    std::vector<std::vector<int>> pre;  // prefix data in insertion order
    std::vector<int> fath;

    int find(const std::vector<int> &d) const {
        for (size_t i = 0; i < pre.size(); i++)
            if (pre[i] == d) return (int)i;
        return -1;
    }
    // schedule.hpp:655-675
    int insert(const std::vector<int> &d) {
        if (d.empty()) return -1;
        int pos = find(d);
        if (pos >= 0) return pos;
        int f = insert(std::vector<int>(d.begin(), d.end() - 1));
        pre.push_back(d);
        fath.push_back(f);
        return (int)pre.size() - 1;
    }
};

}  // namespace

static int check_pattern(const Pattern &p) {
    if (p.n < 3) GAT_FAIL(GAT_EUNSUPPORTED, "patterns need at least 3 vertices (got %d)", p.n);
    if (p.n > 8)
        GAT_FAIL(GAT_ELIMIT, "schedule search is O(n!^2); patterns above 8 vertices are refused");
    uint32_t seen = 1, frontier = 1;  // connectivity: some order must be valid
    while (frontier) {
        uint32_t nxt = 0;
        for (int v = 0; v < p.n; v++)
            if (frontier >> v & 1) nxt |= p.row[v];
        frontier = nxt & ~seen;
        seen |= nxt;
    }
    if (seen != (1u << p.n) - 1) GAT_FAIL(GAT_EINVAL, "pattern is not connected");
    return GAT_OK;
}

int schedule_candidates(const Pattern &p, const ScheduleStats &st, std::vector<ScheduleCandidate> *out) {
    int rc = check_pattern(p);
    if (rc) return rc;
    const std::vector<Perm> aut = automorphisms(p);
    // ---- permutation search (schedule.hpp:864-926): every (permutation,
    // restriction set) at the largest IEP suffix, in scan order ----
    std::vector<ScheduleCandidate> all;
    int best_suffix = -1;
    Perm perm(p.n);
    std::iota(perm.begin(), perm.end(), 0);
    do {
        const Pattern q = p.relabel(perm.data());
        if (!order_connected(q)) continue;
        const int suf = iep_suffix_length(q);
        if (suf < best_suffix) continue;
        if (suf > best_suffix) {
            best_suffix = suf;
            all.clear();
        }
        // Aut(q) = perm o Aut(p) o perm^-1, listed in lexicographic order
        // as a direct enumeration of Aut(q) would produce it.
        std::vector<Perm> autq;
        for (const auto &a : aut) {
            Perm t(p.n);
            for (int i = 0; i < p.n; i++) t[perm[i]] = perm[a[i]];
            autq.push_back(std::move(t));
        }
        std::sort(autq.begin(), autq.end());
        for (const auto &r : restriction_candidates_from(q, autq))
            all.push_back(ScheduleCandidate{perm, r, estimate_restrictions(q, r, st.v_cnt, st.e_cnt, st.tri_cnt)});
    } while (std::next_permutation(perm.begin(), perm.end()));
    if (all.empty()) GAT_FAIL(GAT_EINVAL, "no valid schedule for pattern %s", p.str().c_str());
    // The reference keeps the first candidate of strictly smallest cost, so
    // a stable sort puts its pick first. Candidates that relabel the pattern
    // identically and carry the same restriction set compile to the same
    // plan: only the first of them is kept.
    std::stable_sort(all.begin(), all.end(),
                     [](const ScheduleCandidate &x, const ScheduleCandidate &y) { return x.cost < y.cost; });
    std::set<std::pair<std::vector<uint16_t>, std::vector<Pair>>> seen;
    out->clear();
    for (auto &c : all) {
        const Pattern q = p.relabel(c.perm.data());
        std::vector<Pair> r = c.restrictions;
        std::sort(r.begin(), r.end());
        if (seen.insert({std::vector<uint16_t>(q.row, q.row + q.n), r}).second) out->push_back(std::move(c));
    }
    return GAT_OK;
}

int build_schedule(const Pattern &p, Schedule *out) {
    std::vector<ScheduleCandidate> c;
    int rc = schedule_candidates(p, ScheduleStats{}, &c);
    if (rc) return rc;
    return build_schedule_from(p, c[0], out);
}

int build_schedule_from(const Pattern &p, const ScheduleCandidate &cand, Schedule *out) {
    int rc = check_pattern(p);
    if (rc) return rc;
    Schedule s;
    s.input = p;
    s.n = p.n;
    s.automorphism_count = (int)automorphisms(p).size();
    const Perm &best_perm = cand.perm;
    const std::vector<Pair> &best_r = cand.restrictions;
    for (int i = 0; i < p.n; i++) s.perm[i] = best_perm[i];
    s.best_cost = cand.cost;
    const Pattern q = p.relabel(best_perm.data());
    s.ordered = q;
    s.chosen = best_r;

    // ---- calculate_pattern (schedule.hpp:677-750) ----
    s.iep_suffix = iep_suffix_length(q);
    s.basic = q.n - s.iep_suffix;
    Builder b{s, {}, {}};
    for (int i = 0; i < s.basic; i++) {
        std::vector<int> d;
        for (int j = 0; j < i; j++)
            if (q.has_edge(i, j)) d.push_back(j);
        b.insert(d);
    }
    std::vector<int> vid_prefix;  // iep_vertex_id before sorting
    for (const auto &g : iep_groups(s.iep_suffix)) {
        const int base = q.n - s.iep_suffix;
        for (size_t sg = 0; sg < g.blocks.size(); sg++) {
            std::vector<int> d;
            for (int i = 0; i < base; i++)
                for (int v : g.blocks[sg])
                    if (q.has_edge(i, base + v)) {
                        d.push_back(i);
                        break;
                    }
            const int pid = b.insert(d);
            auto it = std::find(vid_prefix.begin(), vid_prefix.end(), pid);
            const int vid = (int)(it - vid_prefix.begin());
            if (it == vid_prefix.end()) vid_prefix.push_back(pid);
            s.iep_ans_pos.push_back(vid);
            const bool last = sg + 1 == g.blocks.size();
            s.iep_coef.push_back(last ? g.coef : 0);
            s.iep_flag.push_back(last ? 1 : 0);
        }
    }
    if ((int)b.pre.size() > GAT_MAX_PREFIXS)
        GAT_FAIL(GAT_ELIMIT, "schedule needs %zu prefixes (> %d)", b.pre.size(), GAT_MAX_PREFIXS);
    if ((int)s.iep_ans_pos.size() > GAT_MAX_IEP_GROUPS)
        GAT_FAIL(GAT_ELIMIT, "schedule needs %zu IEP subgroups", s.iep_ans_pos.size());

    // ---- sort_prefixs (schedule.hpp:614-652) ----
    const int np = (int)b.pre.size();
    std::vector<int> order(np);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return std::make_pair(b.pre[x].back(), x) < std::make_pair(b.pre[y].back(), y);
    });
    std::vector<int> where(np);
    for (int i = 0; i < np; i++) where[order[i]] = i;
    for (int i = 0; i < np; i++) {
        s.prefixes.push_back(Prefix{b.pre[order[i]]});
        const int f = b.fath[order[i]];
        s.father.push_back(f < 0 ? -1 : where[f]);
    }
    for (int pid : vid_prefix) s.iep_vertex_id.push_back(where[pid]);

    // ---- build_loop_invariant (schedule.hpp:786-818) ----
    s.loop_set_prefix.push_back(-1);
    for (int i = 1; i < s.basic; i++) {
        std::vector<int> d;
        for (int j = 0; j < i; j++)
            if (q.has_edge(j, i)) d.push_back(j);
        int pos = -1;
        for (int k = 0; k < np; k++)
            if (s.prefixes[k].data == d) {
                pos = k;
                break;
            }
        if (pos < 0) GAT_FAIL(GAT_EINVAL, "prefix not found for vertex %d", i);
        s.loop_set_prefix.push_back(pos);
    }
    int cur = 0;
    for (int i = 0; i < s.basic; i++) {
        s.vertex_prefix_start.push_back(cur);
        while (cur < np && s.prefixes[cur].last() == i) cur++;
    }
    s.vertex_prefix_start.push_back(np);

    // ---- build_restrictions (schedule.hpp:820-845) ----
    s.filtered = best_r;
    std::sort(s.filtered.begin(), s.filtered.end(), [](const Pair &x, const Pair &y) {
        if (x.second != y.second) return x.second < y.second;
        return x.first < y.first;
    });
    s.filtered.erase(std::remove_if(s.filtered.begin(), s.filtered.end(),
                                    [&](const Pair &x) { return x.second >= s.basic; }),
                     s.filtered.end());
    // The reference bounds this scan by basic_vertexes rather than by the
    // list length and so can read past the vector; the scan here stops at
    // the end of the list, which is what the loop intends.
    int cs = 0;
    for (int i = 0; i < s.basic; i++) {
        while (cs < (int)s.filtered.size() && s.filtered[cs].second < i) cs++;
        s.restrictions_start.push_back(cs);
    }
    s.restrictions_start.push_back((int)s.filtered.size());

    // ---- build_only_need_size (schedule.hpp:847-862) ----
    std::vector<int> has_child(np, 0);
    for (int i = 0; i < np; i++)
        if (s.father[i] != -1) has_child[s.father[i]] = 1;
    for (int i = 0; i < np; i++) s.size_only.push_back(has_child[i] ? 0 : 1);

    *out = std::move(s);
    return GAT_OK;
}

void export_schedule(const Schedule &s, gat_schedule_data *o) {
    std::memset(o, 0, sizeof(*o));
    o->v_cnt = s.n;
    o->iep_suffix_vertexes = s.iep_suffix;
    o->basic_vertexes = s.basic;
    o->total_prefix_num = (int)s.prefixes.size();
    o->total_restrictions_num = (int)s.filtered.size();
    for (int i = 0; i < GAT_MAX_PREFIXS; i++) o->prefixs_father[i] = -1;
    for (size_t i = 0; i < s.prefixes.size(); i++) {
        o->prefix_depth[i] = (int)s.prefixes[i].data.size();
        for (size_t k = 0; k < s.prefixes[i].data.size(); k++) o->prefix_data[i][k] = s.prefixes[i].data[k];
        o->prefixs_father[i] = s.father[i];
        o->prefixs_size_only[i] = s.size_only[i];
    }
    for (int i = 0; i <= s.basic; i++) {
        o->vertex_prefix_start[i] = s.vertex_prefix_start[i];
        o->restrictions_start[i] = s.restrictions_start[i];
    }
    for (int i = 0; i < s.basic; i++) o->loop_set_prefix_id[i] = s.loop_set_prefix[i];
    for (size_t i = 0; i < s.filtered.size(); i++) {
        o->restrictions[i] = s.filtered[i].first;
        o->restriction_second[i] = s.filtered[i].second;
    }
    o->iep_prefix_num = (int)s.iep_vertex_id.size();
    o->subgroups_num = (int)s.iep_ans_pos.size();
    for (size_t i = 0; i < s.iep_vertex_id.size(); i++) o->iep_vertex_id[i] = s.iep_vertex_id[i];
    for (size_t i = 0; i < s.iep_ans_pos.size(); i++) {
        o->iep_ans_pos[i] = s.iep_ans_pos[i];
        o->iep_coef[i] = s.iep_coef[i];
        o->iep_flag[i] = s.iep_flag[i];
    }
    for (int i = 0; i < s.n; i++) {
        o->perm[i] = s.perm[i];
        o->adj_rows[i] = s.ordered.row[i];
    }
    o->n_all_restrictions = (int)s.chosen.size();
    for (size_t i = 0; i < s.chosen.size(); i++) {
        o->all_restr_first[i] = s.chosen[i].first;
        o->all_restr_second[i] = s.chosen[i].second;
    }
    o->automorphisms = s.automorphism_count;
    o->best_cost = s.best_cost;
}

}  // namespace gat
