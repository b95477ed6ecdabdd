// Device-side data layout shared by the engine's host code and its kernels.
//
// HBM layout of one engine: a single row-major matrix of uint64 words,
// one row per node, `stride` words per row. Rows [0, n_events) hold the
// basic events in event_idx order; gate rows follow layer by layer. Word s
// of a row covers trials 64*s .. 64*s+63 of the current pass.
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

struct EventDesc {       // 16 bytes, one per basic event, event_idx order
    uint32_t ctr0;       // PDAG index + 1       (Philox counter word 0)
    uint32_t ctr1;       // event_idx + 1         (Philox counter word 1)
    uint32_t threshold;  // p * 2^32, truncated, when mode == kEventDraw
    uint32_t mode_tally; // mode | tallied << 2; tallies accumulate at ones[row]
};

struct GateDesc {        // 16 bytes, one per gate, evaluation order
    uint32_t out_row;
    uint32_t arg_begin;  // into the argument-row array
    uint32_t counts;     // n_in | n_pos << 16  (positive inputs come first)
    uint32_t type_k_tally; // connective | k8 << 8 | tallied << 16
};

struct LaunchGeometry {
    uint64_t stride;        // words per row (multiple of chunk_words)
    uint64_t n_bp;          // live words per row; words >= n_bp are padding
    uint32_t chunk_words;   // words per workgroup task (multiple of 512)
    uint32_t n_chunks;      // stride / chunk_words
};

}  // namespace mcs

#include <hip/hip_runtime_api.h>

namespace mcs {

// Launch wrappers (defined in kernels.hip). Grids are n_events x n_chunks and
// n_gates x n_chunks workgroups of 256 lanes; the caller keeps both products
// below 2^32 / 256.
hipError_t launch_sample_events(hipStream_t stream, const EventDesc *events, uint64_t *words,
                                unsigned long long *ones, const LaunchGeometry &geo,
                                uint32_t n_events, uint32_t iteration, uint32_t k0, uint32_t k1);
hipError_t launch_eval_gates(hipStream_t stream, const GateDesc *gates, const uint32_t *arg_rows,
                             uint64_t *words, unsigned long long *ones, const LaunchGeometry &geo,
                             uint32_t n_gates);

}  // namespace mcs
