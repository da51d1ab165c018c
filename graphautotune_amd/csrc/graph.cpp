#include "graph.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <thread>

namespace gat {

void Graph::finish() {
    max_deg = 0;
    for (int32_t v = 0; v < n; v++) max_deg = std::max(max_deg, deg(v));
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn, int64_t grain) {
    if (n <= 0) return;
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n <= grain || hw == 1) {
        fn(0, n);
        return;
    }
    const int64_t parts = std::min<int64_t>(hw, (n + grain - 1) / grain);
    std::vector<std::thread> ts;
    for (int64_t t = 0; t < parts; t++) {
        const int64_t lo = n * t / parts, hi = n * (t + 1) / parts;
        ts.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    for (auto &t : ts) t.join();
}

int build_csr(int32_t n, const std::vector<int32_t> &src, const std::vector<int32_t> &dst, Graph *g) {
    if (n < 0) GAT_FAIL(GAT_EINVAL, "negative vertex count");
    const int64_t ne = (int64_t)src.size();
    if ((int64_t)dst.size() != ne) GAT_FAIL(GAT_EINVAL, "src/dst length mismatch");
    std::vector<int64_t> cnt(size_t(n) + 1, 0);
    for (int64_t i = 0; i < ne; i++) {
        const int32_t u = src[i], v = dst[i];
        if (u < 0 || v < 0 || u >= n || v >= n)
            GAT_FAIL(GAT_EINVAL, "edge (%d,%d) out of range for %d vertices", u, v, n);
        if (u == v) continue;
        cnt[u + 1]++;
        cnt[v + 1]++;
    }
    for (int32_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
    std::vector<int32_t> tmp(size_t(cnt[n]));
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t i = 0; i < ne; i++) {
            const int32_t u = src[i], v = dst[i];
            if (u == v) continue;
            tmp[pos[u]++] = v;
            tmp[pos[v]++] = u;
        }
    }
    std::vector<int64_t> uniq(size_t(n) + 1, 0);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++) {
            auto b = tmp.begin() + cnt[v], e = tmp.begin() + cnt[v + 1];
            std::sort(b, e);
            uniq[v + 1] = std::unique(b, e) - b;
        }
    });
    g->n = n;
    g->off.assign(size_t(n) + 1, 0);
    for (int32_t v = 0; v < n; v++) g->off[v + 1] = g->off[v] + uniq[v + 1];
    g->m = g->off[n];
    g->adj.assign(size_t(g->m), 0);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++)
            std::copy(tmp.begin() + cnt[v], tmp.begin() + cnt[v] + uniq[v + 1], g->adj.begin() + g->off[v]);
    });
    g->finish();
    return GAT_OK;
}

int check_csr(const Graph &g) {
    if ((int64_t)g.off.size() != int64_t(g.n) + 1 || g.off[0] != 0 || g.off[g.n] != g.m ||
        (int64_t)g.adj.size() != g.m)
        GAT_FAIL(GAT_EINVAL, "CSR offsets inconsistent with entry count");
    std::vector<char> bad(1, 0);
    int32_t bad_v = -1;
    for (int32_t v = 0; v < g.n; v++) {
        if (g.off[v + 1] < g.off[v]) GAT_FAIL(GAT_EINVAL, "CSR offsets decrease at vertex %d", v);
        for (int64_t e = g.off[v]; e < g.off[v + 1]; e++) {
            const int32_t w = g.adj[e];
            if (w < 0 || w >= g.n || w == v || (e > g.off[v] && g.adj[e - 1] >= w)) {
                bad_v = v;
                break;
            }
        }
        if (bad_v >= 0) break;
    }
    if (bad_v >= 0)
        GAT_FAIL(GAT_EINVAL, "row %d is not sorted/unique/loop-free or holds an out-of-range id", bad_v);
    // every entry needs its reverse: the engine's orientation (split[]),
    // hub rows and root shards all assume symmetric adjacency
    std::atomic<int64_t> asym{-1};
    parallel_for(g.n, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi && asym.load(std::memory_order_relaxed) < 0; v++)
            for (int64_t e = g.off[v]; e < g.off[v + 1]; e++) {
                const int32_t w = g.adj[e];
                auto b = g.adj.begin() + g.off[w], en = g.adj.begin() + g.off[w + 1];
                if (!std::binary_search(b, en, int32_t(v))) {
                    asym.store(e, std::memory_order_relaxed);
                    break;
                }
            }
    });
    const int64_t e = asym.load();
    if (e >= 0) {
        const int32_t v = int32_t(std::upper_bound(g.off.begin(), g.off.end(), e) - g.off.begin() - 1);
        GAT_FAIL(GAT_EINVAL, "CSR is not symmetric: entry (%d,%d) has no reverse", v, g.adj[e]);
    }
    return GAT_OK;
}

int load_bin(const std::string &path, Graph *g) {
    std::ifstream f(path, std::ios::binary);
    if (!f) GAT_FAIL(GAT_EIO, "cannot open %s", path.c_str());
    f.seekg(0, std::ios::end);
    const int64_t fsize = f.tellg();
    f.seekg(0);
    int32_t n = 0;
    int64_t m = 0;
    f.read(reinterpret_cast<char *>(&n), 4);
    f.read(reinterpret_cast<char *>(&m), 8);
    if (!f || n < 0 || m < 0 || 12 + 8 * (int64_t(n) + 1) + 4 * m != fsize)
        GAT_FAIL(GAT_EIO, "%s is not a CSR .bin (int32 v, int64 e, int64 off[v+1], int32 adj[e])",
                 path.c_str());
    g->n = n;
    g->m = m;
    g->off.resize(size_t(n) + 1);
    g->adj.resize(size_t(m));
    f.read(reinterpret_cast<char *>(g->off.data()), 8 * (int64_t(n) + 1));
    f.read(reinterpret_cast<char *>(g->adj.data()), 4 * m);
    if (!f) GAT_FAIL(GAT_EIO, "short read on %s", path.c_str());
    int rc = check_csr(*g);
    if (rc) return rc;
    g->finish();
    return GAT_OK;
}

int save_bin(const Graph &g, const std::string &path) {
    std::ofstream f(path, std::ios::binary);
    if (!f) GAT_FAIL(GAT_EIO, "cannot create %s", path.c_str());
    f.write(reinterpret_cast<const char *>(&g.n), 4);
    f.write(reinterpret_cast<const char *>(&g.m), 8);
    f.write(reinterpret_cast<const char *>(g.off.data()), 8 * (int64_t(g.n) + 1));
    f.write(reinterpret_cast<const char *>(g.adj.data()), 4 * g.m);
    if (!f) GAT_FAIL(GAT_EIO, "write failed on %s", path.c_str());
    return GAT_OK;
}

int load_txt(const std::string &path, Graph *g) {
    // graph_backend.cuh:23-89: ids are relabelled by sorted value; the
    // declared vertex count may exceed the ids seen (isolated vertices) but
    // not fall short of them.
    std::ifstream f(path);
    if (!f) GAT_FAIL(GAT_EIO, "cannot open %s", path.c_str());
    int64_t n = 0, e = 0;
    if (!(f >> n >> e) || n < 0 || e < 0 || n > INT32_MAX) GAT_FAIL(GAT_EIO, "bad header in %s", path.c_str());
    std::vector<int64_t> us(e), vs(e);
    for (int64_t i = 0; i < e; i++)
        if (!(f >> us[i] >> vs[i])) GAT_FAIL(GAT_EIO, "%s: expected %lld edges, got %lld", path.c_str(),
                                             (long long)e, (long long)i);
    std::vector<int64_t> ids(us);
    ids.insert(ids.end(), vs.begin(), vs.end());
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if ((int64_t)ids.size() > n)
        GAT_FAIL(GAT_EINVAL, "vertex count %lld is less than the %zu ids that appear", (long long)n, ids.size());
    std::vector<int32_t> src(e), dst(e);
    for (int64_t i = 0; i < e; i++) {
        src[i] = int32_t(std::lower_bound(ids.begin(), ids.end(), us[i]) - ids.begin());
        dst[i] = int32_t(std::lower_bound(ids.begin(), ids.end(), vs[i]) - ids.begin());
    }
    return build_csr(int32_t(n), src, dst, g);
}

int gen_er(int32_t n, double p, uint64_t seed, Graph *g) {
    if (n < 0 || !(p >= 0 && p <= 1)) GAT_FAIL(GAT_EINVAL, "ER needs n >= 0 and 0 <= p <= 1");
    std::vector<int32_t> src, dst;
    if (p > 0 && n > 1) {
        // Batagelj & Brandes (2005) geometric skipping over pairs w < v.
        Rng rng(seed);
        const double lp = std::log1p(-p);
        int64_t v = 1, w = -1;
        while (v < n) {
            const double r = rng.uniform();
            w += 1 + (p >= 1 ? 0 : (int64_t)std::floor(std::log1p(-r) / lp));
            while (w >= v && v < n) {
                w -= v;
                v++;
            }
            if (v < n) {
                src.push_back(int32_t(v));
                dst.push_back(int32_t(w));
            }
        }
    }
    return build_csr(n, src, dst, g);
}

namespace {
constexpr int kStreams = 64;  // fixed stream count keeps output thread-count independent

std::vector<Rng> streams(uint64_t seed) {
    std::vector<Rng> out;
    Rng r(seed);
    for (int k = 0; k < kStreams; k++) {
        out.push_back(r);
        r.jump();
    }
    return out;
}
}  // namespace

int gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, Graph *g) {
    if (scale < 1 || scale > 30 || ef < 1) GAT_FAIL(GAT_EINVAL, "R-MAT needs 1 <= scale <= 30, ef >= 1");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1) GAT_FAIL(GAT_EINVAL, "R-MAT probabilities invalid");
    const int32_t n = int32_t(1) << scale;
    const int64_t m = int64_t(ef) << scale;
    std::vector<int32_t> src(m), dst(m);
    auto rs = streams(seed);
    const double ab = a + b, abc = a + b + c;
    parallel_for(kStreams, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; k++) {
            Rng rng = rs[k];
            for (int64_t i = m * k / kStreams; i < m * (k + 1) / kStreams; i++) {
                int32_t u = 0, v = 0;
                for (int l = 0; l < scale; l++) {
                    const double r = rng.uniform();
                    const int bu = r >= ab, bv = (r >= a && r < ab) || r >= abc;
                    u = (u << 1) | bu;
                    v = (v << 1) | bv;
                }
                src[i] = u;
                dst[i] = v;
            }
        }
    }, 1);
    return build_csr(n, src, dst, g);
}

int gen_chung_lu(int32_t n, double mean_deg, double gamma, int32_t max_deg, uint64_t seed, Graph *g) {
    if (n < 2 || !(mean_deg > 0) || !(gamma > 2) || max_deg < 1)
        GAT_FAIL(GAT_EINVAL, "Chung-Lu needs n >= 2, mean > 0, gamma > 2, max_deg >= 1");
    // Expected-degree weights w_i ~ (i+1)^(-1/(gamma-1)), scaled to the
    // requested mean, then capped.
    std::vector<double> w(n);
    double s = 0;
    for (int32_t i = 0; i < n; i++) s += (w[i] = std::pow(double(i) + 1.0, -1.0 / (gamma - 1.0)));
    const double scale = mean_deg * n / s;
    double tot = 0;
    for (auto &x : w) tot += (x = std::min(x * scale, double(max_deg)));
    // Vose alias table
    std::vector<double> prob(n);
    std::vector<int32_t> alias(n, 0);
    {
        std::vector<int32_t> small, large;
        std::vector<double> q(n);
        for (int32_t i = 0; i < n; i++) {
            q[i] = w[i] * n / tot;
            (q[i] < 1.0 ? small : large).push_back(i);
        }
        while (!small.empty() && !large.empty()) {
            const int32_t l = small.back(), h = large.back();
            small.pop_back();
            prob[l] = q[l];
            alias[l] = h;
            q[h] = (q[h] + q[l]) - 1.0;
            if (q[h] < 1.0) {
                large.pop_back();
                small.push_back(h);
            }
        }
        for (int32_t i : large) prob[i] = 1.0;
        for (int32_t i : small) prob[i] = 1.0;
    }
    const int64_t m = (int64_t)std::llround(double(n) * mean_deg / 2.0);
    std::vector<int32_t> src(m), dst(m);
    auto rs = streams(seed);
    parallel_for(kStreams, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; k++) {
            Rng rng = rs[k];
            auto draw = [&]() -> int32_t {
                const int32_t i = int32_t(rng.below(uint64_t(n)));
                return rng.uniform() < prob[i] ? i : alias[i];
            };
            for (int64_t i = m * k / kStreams; i < m * (k + 1) / kStreams; i++) {
                src[i] = draw();
                dst[i] = draw();
            }
        }
    }, 1);
    return build_csr(n, src, dst, g);
}

void relabel_by_degree(const Graph &in, Graph *out, std::vector<int32_t> *old_of_new) {
    const int32_t n = in.n;
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        const int32_t dx = in.deg(x), dy = in.deg(y);
        return dx != dy ? dx > dy : x < y;
    });
    std::vector<int32_t> nw(n);
    for (int32_t i = 0; i < n; i++) nw[order[i]] = i;
    out->n = n;
    out->m = in.m;
    out->off.assign(size_t(n) + 1, 0);
    for (int32_t i = 0; i < n; i++) out->off[i + 1] = out->off[i] + in.deg(order[i]);
    out->adj.assign(size_t(in.m), 0);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            const int32_t v = order[i];
            int64_t o = out->off[i];
            for (int64_t e = in.off[v]; e < in.off[v + 1]; e++) out->adj[o++] = nw[in.adj[e]];
            std::sort(out->adj.begin() + out->off[i], out->adj.begin() + out->off[i + 1]);
        }
    });
    out->finish();
    if (old_of_new) *old_of_new = std::move(order);
}

int64_t count_triangles(const Graph &g) {
    // below[v] = number of neighbours of v with a smaller id (rows are sorted)
    std::vector<int32_t> below(size_t(g.n), 0);
    parallel_for(g.n, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; v++) {
            const auto b = g.adj.begin() + g.off[v], e = g.adj.begin() + g.off[v + 1];
            below[v] = int32_t(std::lower_bound(b, e, int32_t(v)) - b);
        }
    });
    std::atomic<int64_t> next{0}, total{0};
    const int64_t block = 2048;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&] {
        int64_t local = 0;
        for (;;) {
            const int64_t lo = next.fetch_add(block);
            if (lo >= g.n) break;
            const int64_t hi = std::min<int64_t>(g.n, lo + block);
            for (int64_t v = lo; v < hi; v++) {
                const int32_t *a = g.adj.data() + g.off[v];
                const int32_t na = below[v];
                for (int32_t k = 0; k < na; k++) {
                    const int32_t u = a[k];
                    const int32_t *b = g.adj.data() + g.off[u];
                    const int32_t nb = below[u];
                    int32_t i = 0, j = 0;
                    while (i < na && j < nb) {  // common neighbours w < u < v
                        if (a[i] < b[j]) i++;
                        else if (a[i] > b[j]) j++;
                        else { local++; i++; j++; }
                    }
                }
            }
        }
        total += local;
    };
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < hw; t++) ts.emplace_back(work);
    for (auto &t : ts) t.join();
    return total.load();
}

}  // namespace gat
