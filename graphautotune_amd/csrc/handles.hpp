// Definitions of the opaque C-ABI handles, shared by abi.cpp and engine.hip.
#pragma once
#include <memory>
#include <mutex>

#include "graph.hpp"
#include "plan.hpp"

namespace gat {
struct PlanDevice;  // device buffers of a plan (engine.hip)
void destroy_plan_device(PlanDevice *);
}  // namespace gat

struct gat_graph {
    std::shared_ptr<gat::Graph> g;
    std::shared_ptr<gat::Graph> relabeled;  // lazily built degree order
    int64_t triangles = -1;                 // lazily counted
    std::mutex mu;
};

struct gat_plan {
    gat::CompiledPlan cp;
    std::shared_ptr<gat::Graph> graph;  // the graph the engine runs on
    gat::PlanDevice *dev = nullptr;
    ~gat_plan() {
        if (dev) gat::destroy_plan_device(dev);
    }
};
