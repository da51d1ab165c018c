// pm: command-line counter with the reference's interface
// (target/pm.cu:14-90, result files include/utils/utils.hpp:12-52):
//
//   pm <graph> <pattern> [hash] [--mode reference|unique|labeled] [--config k=v,...] [--root DIR]
//
// <graph> is a CSR .bin (graph_backend.cuh:91-99) or a text edge list
// (.txt), resolved against --root (the reference's PROJECT_ROOT,
// include/consts/project_consts.hpp:5-6; default: current directory).
// Prints "Time: <us> us (<s> s)" and "Answer: <count>"; with a hash it
// writes <root>/tuning_results/<hash>/time.txt and count.txt as the
// reference's tuner expects (tuning/tuning/role/driver.py:118-130). The
// default mode is the reference engine's semantics so scripts reading
// "Answer:" see the same number. Exit code 0 on success, 1 on error.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>

#include "gat.h"

static int fail(const char *what) {
    std::fprintf(stderr, "pm: %s: %s\n", what, gat_last_error());
    return 1;
}

static bool set_knob(gat_kernel_config *c, const std::string &k, int v) {
    struct {
        const char *name;
        int32_t *field;
    } knobs[] = {{"block_waves", &c->block_waves},
                 {"waves_per_cu", &c->waves_per_cu},
                 {"queue_chunk", &c->queue_chunk},
                 {"slot_capacity", &c->slot_capacity},
                 {"leaf_mode", &c->leaf_mode},
                 {"intersect_strategy", &c->intersect_strategy},
                 {"gallop_ratio", &c->gallop_ratio},
                 {"search_linear_threshold", &c->search_linear_threshold},
                 {"max_regs", &c->max_regs},
                 {"vertex_store", &c->vertex_store},
                 {"hub_degree", &c->hub_degree}, {"split_us", &c->split_us}};
    for (auto &kn : knobs)
        if (k == kn.name) {
            *kn.field = v;
            return true;
        }
    return false;
}

static void mkdirs(const std::string &p) {
    std::string cur;
    for (size_t i = 0; i <= p.size(); i++) {
        if (i == p.size() || p[i] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (i < p.size()) cur += p[i];
    }
}

int main(int argc, char **argv) {
    std::string graph, pattern, hash, root = ".", mode = "reference";
    gat_kernel_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    int pos = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--mode" && i + 1 < argc) {
            mode = argv[++i];
        } else if (a == "--root" && i + 1 < argc) {
            root = argv[++i];
        } else if (a == "--config" && i + 1 < argc) {
            std::string s = argv[++i];
            size_t st = 0;
            while (st < s.size()) {
                size_t e = s.find(',', st);
                if (e == std::string::npos) e = s.size();
                std::string kv = s.substr(st, e - st);
                size_t eq = kv.find('=');
                if (eq == std::string::npos || !set_knob(&cfg, kv.substr(0, eq), std::atoi(kv.c_str() + eq + 1))) {
                    std::fprintf(stderr, "pm: bad --config entry '%s'\n", kv.c_str());
                    return 1;
                }
                st = e + 1;
            }
        } else if (pos == 0) {
            graph = a, pos++;
        } else if (pos == 1) {
            pattern = a, pos++;
        } else if (pos == 2) {
            hash = a, pos++;
        } else {
            std::fprintf(stderr, "pm: unexpected argument '%s'\n", a.c_str());
            return 1;
        }
    }
    if (pos < 2) {
        std::fprintf(stderr, "Usage: ./pm <graph_name> <pattern_str> [hash_code] [--mode reference|unique|labeled] "
                             "[--config k=v,...] [--root DIR]\n");
        return 1;
    }
    const int m = mode == "unique" ? GAT_MODE_UNIQUE : mode == "labeled" ? GAT_MODE_LABELED : GAT_MODE_REFERENCE;
    if (mode != "unique" && mode != "labeled" && mode != "reference") {
        std::fprintf(stderr, "pm: unknown mode '%s'\n", mode.c_str());
        return 1;
    }
    const std::string path = (graph.size() && graph[0] == '/') ? graph : root + "/" + graph;
    std::printf("Graph Name: %s\n", graph.c_str());
    if (!hash.empty()) std::printf("Hash Code: %s\n", hash.c_str());
    gat_graph *g = nullptr;
    const bool txt = path.size() > 4 && path.substr(path.size() - 4) == ".txt";
    if ((txt ? gat_graph_load_txt(path.c_str(), &g) : gat_graph_load_bin(path.c_str(), &g)) != GAT_OK)
        return fail("cannot load the graph");
    int32_t n = 0, maxd = 0;
    int64_t e = 0;
    gat_graph_info(g, &n, &e, &maxd);
    std::printf("Vertex Count: %d\nEdge Count: %lld\nMax Degree: %d\n", n, (long long)e, maxd);
    gat_plan *plan = nullptr;
    const char *pats[1] = {pattern.c_str()};
    if (gat_plan_create(g, pats, 1, m, &plan) != GAT_OK) return fail("cannot build the schedule");
    uint64_t count = 0;
    gat_count_stats st;
    if (gat_count(plan, &cfg, 0, &count, &st) != GAT_OK) return fail("count failed");
    const double us = st.seconds * 1e6;
    std::printf("Time: %.0f us (%g s)\n", us, st.seconds);
    if (m == GAT_MODE_REFERENCE)
        std::printf("Answer: %lld\n", (long long)count);  // the reference prints a long long
    else
        std::printf("Answer: %llu\n", (unsigned long long)count);
    if (!hash.empty()) {
        const std::string dir = root + "/tuning_results/" + hash;
        mkdirs(dir);
        FILE *ft = std::fopen((dir + "/time.txt").c_str(), "w");
        FILE *fc = std::fopen((dir + "/count.txt").c_str(), "w");
        if (!ft || !fc) {
            std::fprintf(stderr, "pm: cannot write result files in %s: %s\n", dir.c_str(), std::strerror(errno));
            return 1;
        }
        std::fprintf(ft, "%g\n", st.seconds);
        if (m == GAT_MODE_REFERENCE)
            std::fprintf(fc, "%lld\n", (long long)count);
        else
            std::fprintf(fc, "%llu\n", (unsigned long long)count);
        std::fclose(ft);
        std::fclose(fc);
    }
    gat_plan_free(plan);
    gat_graph_free(g);
    return 0;
}
