// Host-side C ABI: errors, schedules, graphs, plans. The GPU entry points
// (gat_count*, gat_copy_probe, gat_device_count) live in engine.hip.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <new>

#include "handles.hpp"

namespace gat {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char *last_error() { return g_err; }

}  // namespace gat

using namespace gat;

#define GUARD_BEGIN try {
#define GUARD_END                                                   \
    }                                                               \
    catch (const std::bad_alloc &) {                                \
        GAT_FAIL(GAT_ENOMEM, "host allocation failed");             \
    }                                                               \
    catch (const std::exception &ex) {                              \
        GAT_FAIL(GAT_EINVAL, "%s", ex.what());                      \
    }

extern "C" {

const char *gat_last_error(void) { return last_error(); }
const char *gat_version(void) { return "gat-mi355x 0.1 (gfx950)"; }

int gat_schedule_build(const char *adj_str, gat_schedule_data *out, size_t cap) {
    if (!adj_str || !out) GAT_FAIL(GAT_EINVAL, "null argument");
    if (cap < sizeof(gat_schedule_data)) GAT_FAIL(GAT_EINVAL, "schedule buffer too small (%zu < %zu)", cap,
                                                  sizeof(gat_schedule_data));
    GUARD_BEGIN
    Pattern p;
    int rc = parse_pattern(adj_str, &p);
    if (rc) return rc;
    Schedule s;
    rc = build_schedule(p, &s);
    if (rc) return rc;
    export_schedule(s, out);
    return GAT_OK;
    GUARD_END
}

int gat_schedule_candidates(const char *adj_str, int32_t v_cnt, int64_t e_cnt, int64_t tri_cnt,
                            gat_schedule_data *out, size_t elem_size, int32_t cap, int32_t *n_out) {
    if (!adj_str || !n_out || cap < 0 || (cap > 0 && !out)) GAT_FAIL(GAT_EINVAL, "bad arguments");
    if (cap > 0 && elem_size < sizeof(gat_schedule_data))
        GAT_FAIL(GAT_EINVAL, "schedule record too small (%zu < %zu)", elem_size, sizeof(gat_schedule_data));
    if (v_cnt <= 0 || e_cnt <= 0 || tri_cnt < 0) GAT_FAIL(GAT_EINVAL, "graph statistics must be positive");
    GUARD_BEGIN
    Pattern p;
    int rc = parse_pattern(adj_str, &p);
    if (rc) return rc;
    std::vector<ScheduleCandidate> c;
    rc = schedule_candidates(p, ScheduleStats{v_cnt, e_cnt, tri_cnt}, &c);
    if (rc) return rc;
    *n_out = (int32_t)c.size();
    for (int32_t i = 0; i < cap && i < (int32_t)c.size(); i++) {
        Schedule s;
        rc = build_schedule_from(p, c[i], &s);
        if (rc) return rc;
        export_schedule(s, reinterpret_cast<gat_schedule_data *>(reinterpret_cast<char *>(out) + i * elem_size));
    }
    return GAT_OK;
    GUARD_END
}

int gat_pattern_automorphisms(const char *adj_str, int64_t *out) {
    if (!adj_str || !out) GAT_FAIL(GAT_EINVAL, "null argument");
    GUARD_BEGIN
    Pattern p;
    int rc = parse_pattern(adj_str, &p);
    if (rc) return rc;
    *out = (int64_t)automorphisms(p).size();
    return GAT_OK;
    GUARD_END
}

static int wrap_graph(Graph &&g, gat_graph **out) {
    auto *h = new gat_graph();
    h->g = std::make_shared<Graph>(std::move(g));
    *out = h;
    return GAT_OK;
}

#define GRAPH_CALL(expr)                  \
    if (!out) GAT_FAIL(GAT_EINVAL, "null out"); \
    GUARD_BEGIN                           \
    Graph g;                              \
    int rc = (expr);                      \
    if (rc) return rc;                    \
    return wrap_graph(std::move(g), out); \
    GUARD_END

int gat_graph_load_bin(const char *path, gat_graph **out) { GRAPH_CALL(load_bin(path ? path : "", &g)) }
int gat_graph_load_txt(const char *path, gat_graph **out) { GRAPH_CALL(load_txt(path ? path : "", &g)) }
int gat_graph_gen_er(int32_t n, double p, uint64_t seed, gat_graph **out) { GRAPH_CALL(gen_er(n, p, seed, &g)) }
int gat_graph_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, gat_graph **out) {
    GRAPH_CALL(gen_rmat(scale, ef, a, b, c, seed, &g))
}
int gat_graph_gen_chung_lu(int32_t n, double mean, double gamma, int32_t maxd, uint64_t seed, gat_graph **out) {
    GRAPH_CALL(gen_chung_lu(n, mean, gamma, maxd, seed, &g))
}

int gat_graph_from_csr(int32_t v_cnt, int64_t e, const int64_t *offsets, const int32_t *adj, gat_graph **out) {
    if (!out || v_cnt < 0 || e < 0 || !offsets || (e > 0 && !adj)) GAT_FAIL(GAT_EINVAL, "bad CSR arguments");
    GUARD_BEGIN
    Graph g;
    g.n = v_cnt;
    g.m = e;
    g.off.assign(offsets, offsets + v_cnt + 1);
    g.adj.assign(adj, adj + e);
    int rc = check_csr(g);  // sorted, unique, loop-free, in range, symmetric
    if (rc) return rc;
    g.finish();
    return wrap_graph(std::move(g), out);
    GUARD_END
}

int gat_graph_from_edges(int32_t v_cnt, int64_t n_edges, const int32_t *src, const int32_t *dst, gat_graph **out) {
    if (!out || n_edges < 0 || (n_edges > 0 && (!src || !dst))) GAT_FAIL(GAT_EINVAL, "bad edge-list arguments");
    GUARD_BEGIN
    std::vector<int32_t> s(src, src + n_edges), d(dst, dst + n_edges);
    Graph g;
    int rc = build_csr(v_cnt, s, d, &g);
    if (rc) return rc;
    return wrap_graph(std::move(g), out);
    GUARD_END
}

int gat_graph_save_bin(const gat_graph *g, const char *path) {
    if (!g || !path) GAT_FAIL(GAT_EINVAL, "null argument");
    GUARD_BEGIN
    return save_bin(*g->g, path);
    GUARD_END
}

int gat_graph_info(const gat_graph *g, int32_t *v_cnt, int64_t *e, int32_t *max_degree) {
    if (!g) GAT_FAIL(GAT_EINVAL, "null graph");
    if (v_cnt) *v_cnt = g->g->n;
    if (e) *e = g->g->m;
    if (max_degree) *max_degree = g->g->max_deg;
    return GAT_OK;
}

int gat_graph_csr(const gat_graph *g, const int64_t **offsets, const int32_t **adj) {
    if (!g) GAT_FAIL(GAT_EINVAL, "null graph");
    if (offsets) *offsets = g->g->off.data();
    if (adj) *adj = g->g->adj.data();
    return GAT_OK;
}

int gat_graph_degree_histogram(const gat_graph *g, int64_t *bins) {
    if (!g || !bins) GAT_FAIL(GAT_EINVAL, "null argument");
    std::memset(bins, 0, sizeof(int64_t) * 32);
    for (int32_t v = 0; v < g->g->n; v++) {
        const uint32_t d = (uint32_t)g->g->deg(v) + 1;
        bins[31 - __builtin_clz(d)]++;
    }
    return GAT_OK;
}

void gat_graph_free(gat_graph *g) { delete g; }

static std::shared_ptr<Graph> relabeled_graph(gat_graph *g) {  // caller holds g->mu
    if (!g->relabeled) {
        auto r = std::make_shared<Graph>();
        relabel_by_degree(*g->g, r.get(), nullptr);
        g->relabeled = r;
    }
    return g->relabeled;
}

static int64_t graph_triangles(gat_graph *g) {  // caller holds g->mu
    if (g->triangles < 0) g->triangles = count_triangles(*relabeled_graph(g));
    return g->triangles;
}

int gat_graph_triangles(gat_graph *g, int64_t *out) {
    if (!g || !out) GAT_FAIL(GAT_EINVAL, "null argument");
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(g->mu);
    *out = graph_triangles(g);
    return GAT_OK;
    GUARD_END
}

int gat_plan_create_ex(gat_graph *g, const char *const *adj_strs, int32_t n_patterns, int32_t mode,
                       const gat_plan_options *opt, gat_plan **out) {
    if (!g || !adj_strs || n_patterns < 1 || !out) GAT_FAIL(GAT_EINVAL, "bad plan arguments");
    GUARD_BEGIN
    std::vector<std::string> adj;
    for (int i = 0; i < n_patterns; i++) {
        if (!adj_strs[i]) GAT_FAIL(GAT_EINVAL, "null pattern %d", i);
        adj.emplace_back(adj_strs[i]);
    }
    PlanOptions po;
    if (opt) {
        po.schedule_policy = opt->schedule_policy;
        po.schedule_rank = opt->schedule_rank;
        if (opt->schedule_ranks) po.schedule_ranks.assign(opt->schedule_ranks, opt->schedule_ranks + n_patterns);
        if (po.schedule_policy == GAT_SCHEDULE_GRAPH && mode != GAT_MODE_REFERENCE) {
            std::lock_guard<std::mutex> lk(g->mu);
            const int64_t tri = graph_triangles(g);
            if (g->g->n <= 0 || g->g->m <= 0)
                GAT_FAIL(GAT_EINVAL, "graph-aware schedules need a graph with edges");
            po.stats = ScheduleStats{g->g->n, g->g->m, std::max<int64_t>(tri, 1)};
        }
    }
    auto *p = new gat_plan();
    int rc = compile_plan(adj, mode, po, &p->cp);
    if (rc) {
        delete p;
        return rc;
    }
    if (p->cp.relabel) {
        std::lock_guard<std::mutex> lk(g->mu);
        p->graph = relabeled_graph(g);
    } else {
        p->graph = g->g;
    }
    *out = p;
    return GAT_OK;
    GUARD_END
}

int gat_plan_create(gat_graph *g, const char *const *adj_strs, int32_t n_patterns, int32_t mode, gat_plan **out) {
    return gat_plan_create_ex(g, adj_strs, n_patterns, mode, nullptr, out);
}

int gat_plan_schedule(const gat_plan *p, int32_t i, gat_schedule_data *out, size_t cap) {
    if (!p || !out || i < 0 || i >= (int)p->cp.patterns.size()) GAT_FAIL(GAT_EINVAL, "bad arguments");
    if (cap < sizeof(gat_schedule_data)) GAT_FAIL(GAT_EINVAL, "schedule buffer too small");
    export_schedule(p->cp.patterns[i].sched, out);
    return GAT_OK;
}

int gat_plan_count_info(const gat_plan *p, int32_t i, gat_pattern_count_info *out, size_t cap) {
    if (!p || !out || i < 0 || i >= (int)p->cp.patterns.size()) GAT_FAIL(GAT_EINVAL, "bad arguments");
    if (cap < sizeof(gat_pattern_count_info)) GAT_FAIL(GAT_EINVAL, "count-info buffer too small");
    std::memset(out, 0, sizeof(*out));
    const PatternInfo &pi = p->cp.patterns[i];
    out->divisor = pi.divisor;
    out->multiplier = pi.multiplier;
    out->relabeled = p->cp.relabel ? 1 : 0;
    const auto &R = pi.used_restrictions;
    if (R.size() > size_t(GAT_MAX_VERTEXES * GAT_MAX_VERTEXES)) GAT_FAIL(GAT_ELIMIT, "too many restrictions");
    out->n_restrictions = (int32_t)R.size();
    for (size_t k = 0; k < R.size(); k++) {
        out->restr_first[k] = R[k].first;
        out->restr_second[k] = R[k].second;
    }
    return GAT_OK;
}

void gat_plan_free(gat_plan *p) { delete p; }

int gat_default_config(gat_kernel_config *out) {
    if (!out) GAT_FAIL(GAT_EINVAL, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->block_waves = 1;
    out->waves_per_cu = 8;
    out->queue_chunk = 16;
    out->slot_capacity = 1024;
    out->leaf_mode = 0;
    out->intersect_strategy = 0;
    out->gallop_ratio = 32;
    out->search_linear_threshold = 0;
    out->max_regs = 0;
    out->hub_degree = 256;
    out->split_us = 2000;
    return GAT_OK;
}

/* Debug helper: text dump of a plan's device program. */
int gat_plan_describe(const gat_plan *p, char *buf, size_t cap) {
    if (!p || !buf || cap == 0) GAT_FAIL(GAT_EINVAL, "bad arguments");
    std::string s = describe_plan(p->cp.dev);
    for (size_t i = 0; i < p->cp.patterns.size(); i++) {
        const auto &pi = p->cp.patterns[i];
        s += "pattern " + std::to_string(i) + " divisor " + std::to_string(pi.divisor) + " multiplier " +
             std::to_string(pi.multiplier) + " restrictions";
        for (auto r : pi.used_restrictions) s += " (" + std::to_string(r.first) + "," + std::to_string(r.second) + ")";
        s += "\n";
    }
    std::snprintf(buf, cap, "%s", s.c_str());
    return (int)s.size();
}

}  // extern "C"
