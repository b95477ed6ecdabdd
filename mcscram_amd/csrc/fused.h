// Whole-tree-in-LDS evaluation of many independent trees in one persistent
// launch (data shared by the host batch engine and its kernel).
//
// A work item is (tree, range of tiles); a tile is W = 2^log2_w consecutive
// words of one pass. A workgroup holding an item keeps, in LDS, the W words
// of every node of its tree: it samples all basic events of the tile, then
// evaluates the gate layers in order, each layer reading only LDS, and adds
// popcounts into per-node LDS counters that are flushed to HBM once per
// item. Nothing of size N_bp is written to HBM.
#pragma once

#include <cstdint>

#include "kernels.h"

namespace mcs {

constexpr uint32_t kEventTallied = 1u << 8;  // EventDesc::mode flag (fused kernel)

struct FusedTree {
    uint64_t n_bp;         // words per pass
    uint32_t n_ev;         // basic events (node ids 0 .. n_ev-1)
    uint32_t n_nodes;      // events + scheduled gates (gates n_ev .. n_nodes-1)
    uint32_t n_layers;     // gate layers
    uint32_t log2_w;       // tile words W = 1 << log2_w (W <= 64)
    uint32_t ev_begin;     // first EventDesc
    uint32_t gate_begin;   // first FusedGate
    uint32_t layer_begin;  // first entry of layer_off (n_layers + 1 entries)
    uint32_t ones_begin;   // first one-bit counter (one per node)
};

constexpr uint32_t kInlineInputs = 6;
constexpr uint32_t kSlotNegated = 1u << 31;

// Input slots are LDS node slots | kSlotNegated for complemented inputs.
// Slots n_nodes and n_nodes + 1 hold an all-zeros and an all-ones word.
struct FusedGate {         // 32 bytes, evaluation order
    uint32_t n_in;
    uint32_t type_k_flags; // connective | k8 << 8 | kGateTallied; NOT and NULL
                           // are sent as NOR and OR of their last input
    // Input slots when n_in <= kInlineInputs, unused ones padded with the
    // identity of the fold; otherwise in[0] is the offset of the slots in
    // the args array.
    uint32_t in[kInlineInputs];
};

struct WorkItem {
    uint32_t tree;
    uint32_t first_tile;
    uint32_t n_tiles;
    uint32_t reserved;
};

struct FusedLaunch {
    const FusedTree *trees;
    const EventDesc *events;
    const FusedGate *gates;
    const uint32_t *layer_off;
    const uint32_t *args;
    const WorkItem *items;
    unsigned long long *ones;
    uint32_t *state;       // [0] work-queue head, [1] iteration of the pass
    uint32_t n_items;
    uint32_t key0, key1;
};

}  // namespace mcs

#include <hip/hip_runtime_api.h>

namespace mcs {

// LDS bytes a tree needs at tile width W: its nodes and the two identity
// words, W words each, and one 32-bit counter per node.
inline uint64_t fused_lds_bytes(uint32_t n_nodes, uint32_t w) {
    return uint64_t(n_nodes + 2u) * w * 8u + uint64_t(n_nodes) * 4u;
}

hipError_t fused_max_workgroups_per_cu(uint32_t lanes, uint32_t lds_bytes, int *out);
// Starts one pass: queue head = 0, iteration += 1 (a kernel, so the pair can
// be captured in a graph).
hipError_t launch_fused_begin_pass(hipStream_t stream, uint32_t *state);
hipError_t launch_fused(hipStream_t stream, const FusedLaunch &l, uint32_t grid, uint32_t lanes,
                        uint32_t lds_bytes);

// Deepest tree (gate layers of the schedule) the kernel stages in LDS.
constexpr uint32_t kMaxFusedLayers = 1023;
// Largest dynamic LDS a workgroup may ask for on a CDNA4 CU: 160 KiB less
// the kernel's static LDS (queue slot and layer offsets).
constexpr uint32_t kMaxFusedLds = (160u << 10) - 4u * (kMaxFusedLayers + 1u) - 64u;

}  // namespace mcs
