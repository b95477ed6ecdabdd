// Host-side graph compilation: flattened PDAG -> layered evaluation plan.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mcscram_amd.h"

namespace mcs {

struct PlannedGate {
    int32_t pdag_index;
    uint8_t type;
    int32_t k;
    int32_t layer;                 // node depth (>= 1)
    std::vector<int32_t> pos_in;   // input node ids (see Plan::node_of_*)
    std::vector<int32_t> neg_in;
};

// Evaluation plan. Node ids: basic events are 0..n_events-1 in event_idx
// order; gates follow in evaluation order (layer by layer).
struct Plan {
    std::vector<int32_t> event_pdag_index;  // by event_idx
    std::vector<double> event_p;            // by event_idx
    std::vector<PlannedGate> gates;         // evaluation order
    std::vector<uint32_t> layer_begin;      // gate offsets per layer (layers 1..L), size L+1
    std::vector<int32_t> node_pdag_index;   // by node id
    int32_t root_node = -1;

    int32_t n_events() const { return int32_t(event_pdag_index.size()); }
    int32_t n_gates() const { return int32_t(gates.size()); }
    int32_t n_nodes() const { return n_events() + n_gates(); }
    int32_t n_gate_layers() const { return int32_t(layer_begin.size()) - 1; }
};

// Builds the plan the reference's layer_manager builds:
//  * depth per node (variables 0, gate = 1 + max child depth;
//    src/preprocessor.cc:130-176);
//  * the DFS gather from the root (layer_manager.tpp:33-60) followed by the
//    two unstable std::sort calls (tpp:79-82 by depth; tpp:92-115 variables
//    first, gates by connective); event_idx = position in layer 0;
//  * per-gate input order: positive then negated, variables before gates
//    within each (kernel_builder.h:298-426).
// Returns false with a message on malformed input.
bool build_plan(const mcs_pdag &pdag, Plan &plan, std::string &error);

// One gate evaluation of the execution schedule. out is a node id of the
// plan, or an id >= plan.n_nodes() for an auxiliary partial fold.
struct ExecGate {
    int32_t out;
    uint8_t type;
    int32_t k;
    std::vector<int32_t> pos_in, neg_in;
    // Gates evaluated inside this gate's task (see fuse_single_consumer):
    // each appears once among pos_in / neg_in by its node id, and its own
    // inputs are rows.
    std::vector<ExecGate> inl;
};

// Execution schedule: the plan's gates in dependency layers, with every
// AND/OR/XOR/NAND/NOR gate of more than max_fan_in inputs split into a tree
// of partial folds of the same bitwise operation (AND for NAND, OR for NOR),
// the gate itself folding the partials. Bitwise AND, OR and XOR are
// associative and commutative, so every word is unchanged; only the launch
// that evaluates each piece moves. Inputs are grouped in order of depth so
// partials over shallow inputs run in early layers. Event order and tallies
// are those of the plan.
struct Schedule {
    std::vector<std::vector<ExecGate>> layers;  // layers[0] = first gate layer
    int32_t n_aux = 0;
};

Schedule build_schedule(const Plan &plan, uint32_t max_fan_in);

// Layer fusion for the layered kernel. A gate of the plan whose word is
// read by exactly one other gate (its consumer), that is not the root and
// not an ATLEAST gate (whose bit-sliced counter would crowd the consumer's
// registers), is evaluated inside its consumer's task, so its word never
// reaches HBM:
// it is computed in registers and folded in as one more input, and its tally
// is counted there. One level: a gate that holds fused gates is not fused
// itself. Consumers are AND/OR/XOR/NAND/NOR (any order of inputs gives the
// same word) or ATLEAST with at most 15 inputs; a consumer task reads at most
// max_rows rows in all. Partial folds of split wide gates keep their own
// tasks. Afterwards the layers are rebuilt from the fused gates' inputs, so
// a consumer can move to an earlier launch.
void fuse_single_consumer(Schedule &sch, const Plan &plan, uint32_t max_rows);

}  // namespace mcs
