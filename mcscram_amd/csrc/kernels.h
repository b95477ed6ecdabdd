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

#include <hip/hip_runtime_api.h>

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
constexpr uint32_t kGateTallied = 1u << 16;     // popcount into ones[out_row]
constexpr uint32_t kGateNoStore = 1u << 17;     // tally only (basic-event tallies)
constexpr uint32_t kGateForceStore = 1u << 18;  // store the words even when constant
constexpr uint32_t kGateHasFused = 1u << 19;    // some argument is a fused gate (kInlineArg)

// Chunk summaries: one byte per (row, chunk) says whether every live word of
// the row's chunk is zero, every one is all ones, or neither. Every task
// that produces a chunk writes its summary; a gate task reads its inputs'
// summaries first, so a constant input is never loaded (it is folded in as
// a constant, or decides the gate outright: AND with a zero chunk), and a
// gate chunk whose summary is constant is not stored at all. Rows of one
// pass are summarised in that pass, so a summary always describes the words
// (stored or implied) of the pass that reads it.
enum : uint32_t { kChunkZero = 0, kChunkOnes = 1, kChunkMixed = 2 };

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
    uint8_t *state;                // chunk summaries [row][state_stride] (null: none,
                                   // every chunk is stored and loaded)
    uint32_t state_stride;         // bytes per row: n_chunks rounded up to 4
    // Several parameter samples of one tree per launch (uncertainty
    // analysis). Sample s owns the words at words + s * sample_words, the
    // summaries at state + s * sample_state and the tally counters at
    // ones + s * sample_ones. The gate tasks run for samples 0 .. n_samples-1
    // (sample-major, then chunk-major). With samp_events > 0 the sampler
    // range indexes (sample, event, chunk) triples, sample-major, with
    // samp_events events per sample: sample s reads event descriptors
    // events + s * samp_events and draws with key0 = keys0[s] (key0 when
    // keys0 is null). n_samples <= 1 and samp_events = 0 is the one-sample
    // launch of the single-tree engine.
    uint32_t n_samples;
    uint32_t samp_events;
    uint64_t sample_words;
    uint64_t sample_state;
    uint64_t sample_ones;
    const uint32_t *keys0;
};

// Bytes per row of the chunk-summary array.
inline uint32_t state_stride_for(uint32_t n_chunks) { return (n_chunks + 3u) & ~3u; }

// Workgroups of a pass_layer launch are dealt to the XCDs round-robin:
// workgroup b runs on XCD b % kXcds as that XCD's (b / kXcds)-th workgroup.
// XCD x owns the x-th eighth of the launch's gate tasks (chunk-major, so the
// gates it evaluates together read the same chunk of their input rows
// through its own L2) and the x-th eighth of its sampler tasks.
constexpr uint32_t kXcds = 8;

struct XcdShare {
    uint64_t g_lo, s_lo;  // first gate task, first sampler task
    uint64_t ngx, tx;     // gate tasks, all tasks (gate + sampler)
};

__host__ __device__ inline XcdShare xcd_share(uint64_t ng, uint64_t ns, uint32_t x) {
    XcdShare s;
    s.g_lo = ng * x / kXcds;
    s.s_lo = ns * x / kXcds;
    s.ngx = ng * (x + 1) / kXcds - s.g_lo;
    s.tx = s.ngx + (ns * (x + 1) / kXcds - s.s_lo);
    return s;
}

// The XCD interleaves its gate and sampler tasks evenly, without a division
// on the device: with frac = ceil(ngx * 2^32 / tx) (at most 2^32),
// gates_before(j) = min(ngx, floor(j * frac / 2^32)) rises by 0 or 1 per
// workgroup, from 0 at j = 0 to ngx at j = tx. Workgroup j runs gate task
// gates_before(j) if gates_before(j + 1) > gates_before(j), else sampler
// task j - gates_before(j); both run through their whole range once.
// j * frac stays below 2^64 because a launch has fewer than 2^24 workgroups.
inline uint64_t interleave_frac(uint64_t ngx, uint64_t tx) {
    if (tx == 0) return 0;
    const unsigned __int128 n = static_cast<unsigned __int128>(ngx) << 32;
    return static_cast<uint64_t>((n + tx - 1) / tx);
}

__host__ __device__ inline uint64_t gates_before(uint64_t j, uint64_t ngx, uint64_t frac) {
    const uint64_t b = (j * frac) >> 32;
    return b < ngx ? b : ngx;
}

// High 64 bits of a 64 x 64-bit product, from 32-bit halves (on wave-uniform
// operands the compiler keeps it on the scalar unit).
__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    const uint64_t a0 = uint32_t(a), a1 = a >> 32, b0 = uint32_t(b), b1 = b >> 32;
    const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const uint64_t mid = (p00 >> 32) + uint32_t(p01) + uint32_t(p10);
    return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// floor(n / d) from the reciprocal m = floor((2^64 - 1) / d), d >= 1: the
// estimate floor(n m / 2^64) is floor(n / d) or one less (m > 2^64/d - 1 and
// n < 2^64), so one correction step makes it exact.
inline uint64_t reciprocal64(uint64_t d) { return d ? ~uint64_t(0) / d : 0; }

__host__ __device__ inline uint64_t div_by_reciprocal(uint64_t n, uint64_t d, uint64_t m) {
    uint64_t q = mulhi64(n, m);
    if (n - q * d >= d) ++q;
    return q;
}

// Per-launch constants computed on the host, so the kernel's task decode
// does no division.
struct SplitFracs {
    uint64_t frac[kXcds];  // interleave_frac of each XCD's share
    uint64_t inv_gates;    // reciprocal64(n_gates)
    uint64_t inv_chunks;   // reciprocal64(n_chunks)
    uint64_t inv_sample_gate_tasks;  // reciprocal64(n_gates * n_chunks)
    uint64_t inv_sample_samp_tasks;  // reciprocal64(samp_events * n_chunks)
};

// Gate tasks of a launch: one per (sample, gate, chunk).
__host__ __device__ inline uint64_t gate_tasks(const PassLaunch &l, uint32_t n_chunks) {
    return uint64_t(l.n_gates) * n_chunks * (l.n_samples > 1 ? l.n_samples : 1u);
}

// Blocks of 256 lanes; the caller checks grid_blocks(...) * 256 < 2^32.
uint64_t pass_grid_blocks(const PassLaunch &l, const LaunchGeometry &geo);
hipError_t launch_pass(hipStream_t stream, const PassLaunch &l, uint64_t *words,
                       unsigned long long *ones, const LaunchGeometry &geo);

}  // namespace mcs
