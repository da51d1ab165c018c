// Pattern graphs and the Schedule builder.
//
// The Schedule is the reference's matching plan (include/core/schedule.hpp):
// it picks a vertex order (permutation) that maximises the inclusion-exclusion
// (IEP) suffix, a symmetry-breaking restriction set, and the prefix trie of
// neighbourhood intersections. Every decision here must equal the
// reference's, because the reference-mode count depends on the exact pick:
// permutations are scanned in next_permutation order, ties on the cost
// estimate keep the first candidate, and restriction sets are produced in the
// reference's depth-first order.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace gat {

constexpr int kMaxV = GAT_MAX_VERTEXES;

// Adjacency as one bitmask row per vertex (n <= 10).
struct Pattern {
    int n = 0;
    uint16_t row[kMaxV] = {};

    bool has_edge(int a, int b) const { return (row[a] >> b) & 1u; }
    void add_edge(int a, int b) {
        row[a] |= uint16_t(1u << b);
        row[b] |= uint16_t(1u << a);
    }
    int edge_count() const;
    int max_degree() const;
    // Vertex i of this pattern becomes vertex perm[i] of the result
    // (Core::get_permutated_pattern, pattern.hpp:74-82).
    Pattern relabel(const int *perm) const;
    bool operator==(const Pattern &o) const;
    std::string str() const;
};

// Parses the n*n row-major '0'/'1' adjacency string; an edge exists when
// either (i,j) or (j,i) is '1' and i != j (pattern.hpp:22-39). Unlike the
// reference, which maps any other character to "no edge" and leaves v_cnt
// at 0 for non-square lengths, malformed strings are rejected.
int parse_pattern(const std::string &s, Pattern *out);

// Every vertex after the first has an earlier neighbour (pattern.hpp:85-101).
bool order_connected(const Pattern &p);

// Length of the trailing run of vertices with no edge to a later vertex,
// capped at n-2 (schedule.hpp:66-83).
int iep_suffix_length(const Pattern &p);

// Automorphisms in lexicographic permutation order (schedule.hpp:173-187).
std::vector<std::vector<int>> automorphisms(const Pattern &p);

using Pair = std::pair<int, int>;

// Number of permutations `data` of 0..n-1 with data[a] < data[b] for every
// (a,b) in pairs (schedule.hpp:126-146 counts the same by brute force).
// Computed here by a DP over vertex subsets.
int64_t linear_extensions(int n, const std::vector<Pair> &pairs);

// Restriction sets that break all symmetry (schedule.hpp:239-360), in the
// reference's order.
std::vector<std::vector<Pair>> restriction_candidates(const Pattern &p);

// Graph-independent cost estimate (schedule.hpp:362-437).
double estimate_restrictions(const Pattern &p, const std::vector<Pair> &restrictions,
                             int v_cnt, int64_t e_cnt, int64_t tri_cnt);

// (#even-edge, #odd-edge) connected labelled graphs on k vertices,
// even-odd = (-1)^(k-1) (k-1)!  (schedule.hpp:149-170 enumerates them).
int iep_block_coefficient(int k);

struct IepGroup {
    std::vector<std::vector<int>> blocks;  // suffix-vertex ids per block
    int coef;
};
// Set partitions of the suffix in the reference's recursion order
// (schedule.hpp:526-572).
std::vector<IepGroup> iep_groups(int suffix);

struct Prefix {
    std::vector<int> data;  // pattern vertices whose neighbourhoods intersect
    int last() const { return data.back(); }
};

struct Schedule {
    Pattern input;    // as given
    Pattern ordered;  // relabelled by perm
    int perm[kMaxV] = {};
    int n = 0;
    int iep_suffix = 0;
    int basic = 0;
    std::vector<Prefix> prefixes;  // sorted by (last vertex, insertion order)
    std::vector<int> father;       // -1 when none
    std::vector<int> size_only;
    std::vector<int> loop_set_prefix;      // [basic], entry 0 = -1
    std::vector<int> vertex_prefix_start;  // [basic + 1]
    std::vector<Pair> chosen;              // best restriction set, unfiltered
    std::vector<Pair> filtered;            // sorted by (second, first), second < basic
    std::vector<int> restrictions_start;   // [basic + 1]
    std::vector<int> iep_vertex_id, iep_ans_pos, iep_coef, iep_flag;
    int automorphism_count = 0;
    double best_cost = 0;
};

// Graph statistics of the cost estimate. The reference's Schedule always
// passes these fixed numbers (schedule.hpp:909-910).
struct ScheduleStats {
    int v_cnt = 5000;
    int64_t e_cnt = 1000000;
    int64_t tri_cnt = 10000;
};

struct ScheduleCandidate {
    std::vector<int> perm;           // vertex i of the input becomes perm[i]
    std::vector<Pair> restrictions;  // on the relabelled pattern
    double cost = 0;                 // estimate_restrictions under the stats
};

// Every (permutation, restriction set) the reference's search considers at
// the largest IEP suffix (schedule.hpp:878-926), ranked by the estimate under
// `st` (stable, so with the fixed stats entry 0 is the reference's pick), with
// duplicates that would compile to the same plan removed.
int schedule_candidates(const Pattern &p, const ScheduleStats &st, std::vector<ScheduleCandidate> *out);

// Core::Schedule(const Pattern&) (schedule.hpp:864-958).
int build_schedule(const Pattern &p, Schedule *out);
// The rest of the Schedule constructor (schedule.hpp:928-958) for a given
// permutation and restriction set.
int build_schedule_from(const Pattern &p, const ScheduleCandidate &c, Schedule *out);

void export_schedule(const Schedule &s, gat_schedule_data *out);

}  // namespace gat
