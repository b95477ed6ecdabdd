// Device-side data layout shared by the engine's host code and its kernels.
//
// HBM layout of one engine: a single row-major matrix of uint64 words,
// `stride` words per row. Word s of a row covers trials 64*s .. 64*s+63 of
// one pass. Rows:
//   [0, n_ev)              basic events of buffer 0, in event_idx order
//   [n_ev, 2 n_ev)         basic events of buffer 1 (when double-buffered)
//   [n_buf n_ev, ...)      gates, layer by layer
// Pass i evaluates its gates from buffer (i & 1) while the same launches
// sample pass i + 1 into the other buffer.
#pragma once

#include <cstdint>

namespace mcs {

// Philox4x32-10 multipliers and Weyl key increments (Salmon et al., SC'11).
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// Sampler modes: a threshold of 0 or >= 2^32 decides every bit without
// drawing (the drawn value is always < 2^32).
enum : uint32_t { kEventDraw = 0, kEventNever = 1, kEventAlways = 2 };

struct EventDesc {     // 16 bytes, one per basic event, event_idx order
    uint32_t ctr0;     // PDAG index + 1       (Philox counter word 0)
    uint32_t ctr1;     // event_idx + 1         (Philox counter word 1)
    uint32_t threshold;// p * 2^32, truncated, when mode == kEventDraw
    uint32_t mode;     // kEventDraw / kEventNever / kEventAlways in bits 0-1
};

// Descriptor of one basic event: threshold uint64(p * 2^32), truncated
// (src/mc/event/node.h:371-376); 0 never draws true and 2^32 always does.
inline EventDesc make_event_desc(int32_t pdag_index, uint32_t event_idx, double p) {
    const uint64_t thr = static_cast<uint64_t>(p * 4294967296.0);
    EventDesc d{};
    d.ctr0 = uint32_t(pdag_index) + 1u;
    d.ctr1 = event_idx + 1u;
    d.threshold = uint32_t(thr < 0xFFFFFFFFull ? thr : 0xFFFFFFFFull);
    d.mode = thr == 0 ? kEventNever : (thr >= (uint64_t(1) << 32) ? kEventAlways : kEventDraw);
    return d;
}

// Gate flags (GateDesc::type_k_flags bits 16..).
constexpr uint32_t kGateTallied = 1u << 16;  // popcount into ones[out_row]
constexpr uint32_t kGateNoStore = 1u << 17;  // tally only (basic-event tallies)

// An argument entry with this bit set is not a row: the low bits index the
// descriptor of a fused gate (in PassLaunch::tree_gates) whose word the task
// computes in registers.
constexpr uint32_t kInlineArg = 1u << 31;

struct GateDesc {          // 16 bytes, one per gate, evaluation order
    uint32_t out_row;      // row written; also the tally slot
    uint32_t arg_begin;    // into the argument-row array
    uint32_t counts;       // n_in | n_pos << 16  (positive inputs come first)
    uint32_t type_k_flags; // connective | k8 << 8 | flags
};

struct LaunchGeometry {
    uint64_t stride;        // words per row (multiple of chunk_words)
    uint64_t n_bp;          // live words per row; words >= n_bp are padding
    uint32_t chunk_words;   // words per workgroup task (512, 1024 or 2048)
    uint32_t n_chunks;      // stride / chunk_words
};

// One launch of the pass kernel: the gates of one layer over every chunk,
// plus a contiguous range of sampler tasks (event-major (event, chunk)
// pairs) for the pass whose basic events go to rows samp_row0 + e.
struct PassLaunch {
    const GateDesc *gates;
    const GateDesc *tree_gates;    // base of kInlineArg references
    const uint32_t *arg_rows;
    uint32_t n_gates;
    uint32_t samp_iteration;       // iteration sampled (when iteration_ptr is null)
    const uint32_t *iteration_ptr; // else *iteration_ptr + samp_iteration
    const EventDesc *events;
    uint64_t samp_begin, samp_end;
    uint32_t samp_row0;
    uint32_t key0, key1;
};

}  // namespace mcs

#include <hip/hip_runtime_api.h>

namespace mcs {

// Blocks of 256 lanes; the caller checks grid_blocks(...) * 256 < 2^32.
uint64_t pass_grid_blocks(const PassLaunch &l, const LaunchGeometry &geo);
hipError_t launch_pass(hipStream_t stream, const PassLaunch &l, uint64_t *words,
                       unsigned long long *ones, const LaunchGeometry &geo);

}  // namespace mcs
