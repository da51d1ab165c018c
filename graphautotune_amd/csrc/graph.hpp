// CSR data graph: loaders, writer, seeded generators and relabelling.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace gat {

struct DeviceGraph;  // engine-side copy, defined in the HIP translation unit

// Same layout as Infra::GlobalMemoryGraph (graph_backend.cuh:16-21): int64
// offsets[n+1] into an int32 adjacency of m directed entries, each row sorted
// ascending without duplicates or self-loops.
struct Graph {
    int32_t n = 0;
    int64_t m = 0;  // directed entries (= 2 * undirected edges)
    std::vector<int64_t> off;
    std::vector<int32_t> adj;
    int32_t max_deg = 0;
    int32_t max_out = -1;       // max #neighbours below the vertex (computed on demand)
    std::shared_ptr<void> dev;  // device copy, created by the engine on first use

    int32_t deg(int32_t v) const { return int32_t(off[v + 1] - off[v]); }
    void finish();  // recompute max_deg
};

// Runs fn(lo, hi) over [0, n) split across hardware threads.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn, int64_t grain = 4096);

// Builds a CSR from undirected edges: drops self-loops, symmetrises, sorts
// rows and removes duplicates (graph_backend.cuh:51-88 does the same with
// per-row std::vector).
int build_csr(int32_t n, const std::vector<int32_t> &src, const std::vector<int32_t> &dst, Graph *g);
int check_csr(const Graph &g);

int load_bin(const std::string &path, Graph *g);
int load_txt(const std::string &path, Graph *g);
int save_bin(const Graph &g, const std::string &path);

int gen_er(int32_t n, double p, uint64_t seed, Graph *g);
int gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, Graph *g);
int gen_chung_lu(int32_t n, double mean_deg, double gamma, int32_t max_deg, uint64_t seed, Graph *g);

// Relabels so that new id order is descending degree, ties by ascending old
// id. old_of_new may be null.
void relabel_by_degree(const Graph &in, Graph *out, std::vector<int32_t> *old_of_new);

// Number of triangles w < u < v: for each edge (u < v), the common
// neighbours below u, by merging the two rows' prefixes of smaller ids
// (cheap on a degree-ordered graph, where those prefixes are short).
int64_t count_triangles(const Graph &g);

}  // namespace gat
