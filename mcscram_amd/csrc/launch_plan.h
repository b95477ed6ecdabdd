// Gate descriptors and launch split of one tree for the layered kernel
// (pass_layer), shared by the single-tree engine and the batch program.
#pragma once

#include <cstdint>
#include <vector>

#include "kernels.h"
#include "layering.h"

namespace mcs {

// Where a tree's rows live in a words matrix.
struct RowMap {
    uint32_t be_row0[2];        // first row of event buffer 0 / 1
    uint32_t gate_row0;         // row of plan node n_ev (gates and partials follow)
    uint32_t event_tally_slot0; // tally slot of basic event e = event_tally_slot0 + e
};

struct TreeLaunchPlan {
    // Task descriptors of all launches in order, then the descriptors of
    // fused gates (from launch_begin.back()), which tasks reference by
    // kInlineArg | index among their arguments.
    std::vector<GateDesc> gates;
    std::vector<uint32_t> args[2];       // argument rows with event buffer 0 / 1
    std::vector<uint32_t> launch_begin;  // first gate of each launch (size L + 1)
    std::vector<double> launch_cost;     // words moved per word of pass (size L)
};

// One launch per schedule layer; the first also carries tally-only tasks for
// tallied basic events (their slot is event_tally_slot0 + e).
void build_launch_plan(const Plan &plan, const Schedule &sch, const std::vector<uint8_t> &tallied,
                       const RowMap &rows, TreeLaunchPlan &out);

// Boundaries of `total` sampler tasks split over launches in proportion to
// their cost (size cost.size() + 1, last = total).
std::vector<uint64_t> split_in_proportion(const std::vector<double> &cost, uint64_t total);

}  // namespace mcs
