// Compiles Schedules into the device program of plan_dev.h.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "pattern.hpp"
#include "plan_dev.h"

namespace gat {

struct PatternInfo {
    std::string adj;
    Schedule sched;
    int mode = GAT_MODE_UNIQUE;
    // The engine's raw total equals divisor * (requested count) before the
    // multiplier: unique = raw / divisor, labeled = unique * |Aut|.
    uint64_t divisor = 1;
    uint64_t multiplier = 1;
    std::vector<Pair> used_restrictions;  // restrictions the engine applies
};

struct CompiledPlan {
    GatPlanDev dev;
    std::vector<PatternInfo> patterns;
    bool relabel = false;  // run on the degree-relabelled graph
    // Bitmap-eligible plan: every pattern is a clique K_k with 3 <= k <= 6,
    // counted in unique or labeled mode. Bit k is set for each K_k present;
    // clique_pattern[k] = index of that pattern (first occurrence).
    uint32_t clique_ks = 0;
    int clique_pattern[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
};

// For unique counting: picks the restriction set to apply and the constant
// factor c such that raw count = c * #instances, by checking that every
// ordering of an instance's vertex ids admits the same number of
// automorphic embeddings. Returns c (>= 1).
uint64_t unique_restrictions(const Schedule &s, std::vector<Pair> *used);

// How each pattern's Schedule is chosen. REFERENCE: the reference's pick
// (the only choice whose reference-mode count is defined). GRAPH and RANKED
// (unique and labeled modes): the candidates ranked by the cost estimate
// under `stats` (the graph's own for GRAPH, the reference's fixed numbers for
// RANKED), taking entry `rank` (clamped to the last one).
struct PlanOptions {
    int schedule_policy = GAT_SCHEDULE_REFERENCE;
    int schedule_rank = 0;
    std::vector<int> schedule_ranks;  // per pattern; empty = schedule_rank for all
    ScheduleStats stats;
};

// mode: GAT_MODE_UNIQUE / GAT_MODE_REFERENCE / GAT_MODE_LABELED
int compile_plan(const std::vector<std::string> &adj, int mode, const PlanOptions &opt, CompiledPlan *out);

std::string describe_plan(const GatPlanDev &d);

}  // namespace gat
