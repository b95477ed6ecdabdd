// HIP kernel of the Monte Carlo fault-tree engine, written for CDNA4 (gfx950).
//
// One kernel, pass_layer, does both kinds of work of a pass; each launch
// mixes them so the VALU-bound sampler and the HBM-bound gate evaluation
// share every CU:
//   gate tasks     one gate of the current layer over a chunk of words:
//                  16-byte loads of each input row, bitwise fold, bit-sliced
//                  counter for ATLEAST, popcount tally fused at production.
//                  A task also computes, in registers, the gates fused into
//                  it (gates read by this gate alone): their words never
//                  reach HBM and their tallies are counted here.
//   sampler tasks  Philox4x32-10 Bernoulli words of one basic event over a
//                  chunk of words; lanes own consecutive words so every
//                  per-event value sits in SGPRs and stores are coalesced.
//
// Reference semantics restated here:
//   counter (pdag_idx+1, event_idx+1, sample_idx+1, ((iteration+1)<<6)+g),
//   bit 4g+j = (y_j < threshold)       src/mc/kernel/basic_event.h:112-142,
//                                      src/mc/prng/philox128.h:8-88
//   gate folds and final inversion     src/mc/kernel/gate.h:262-264, 299-362
//   ATLEAST: 8-bit count >= 8-bit k    src/mc/kernel/gate.h:506-555
//   ones += popcount(word)             src/mc/kernel/tally.h:322-360
#include <hip/hip_runtime.h>

#include "device_ops.h"
#include "kernels.h"

namespace mcs {
namespace {

constexpr uint32_t kLanes = 256;

// Row words are streamed: a row is read by a handful of gates spread over the
// whole pass, long after it has left the caches, so loads and stores are
// non-temporal (measured 2 % faster per pass on the large tree).
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void store_pair(uint64_t *p, uint64_t x, uint64_t y) {
    u64x2 v;
    v.x = x;
    v.y = y;
    __builtin_nontemporal_store(v, reinterpret_cast<u64x2 *>(p));
}

// One 64-bit atomic per wave: the whole wave takes part (no lane is
// inactive here), lane 0 adds the wave's total.
__device__ __forceinline__ void tally_wave(unsigned long long *ones, uint32_t slot, uint32_t pc) {
    pc = wave_sum(pc);
    if ((threadIdx.x & 63u) == 0 && pc) atomicAdd(ones + slot, (unsigned long long)pc);
}

// Scalar loads: argument rows and chunk summaries are written before the
// launch that reads them (summaries of a launch's own outputs are never read
// by it), so they are read through the constant address space, which keeps
// the loads on the scalar unit.
typedef const __attribute__((address_space(4))) uint32_t const_u32;

__device__ __forceinline__ uint32_t sload(const uint32_t *p, uint64_t i) {
    return ((const_u32 *)p)[i];
}

// Descriptors are written before the launch too: scalar loads, one dword
// at a time (both are four dwords).
__device__ __forceinline__ GateDesc gate_desc(const GateDesc *p, uint64_t i) {
    const uint32_t *q = reinterpret_cast<const uint32_t *>(p + i);
    return GateDesc{sload(q, 0), sload(q, 1), sload(q, 2), sload(q, 3)};
}

__device__ __forceinline__ EventDesc event_desc(const EventDesc *p, uint64_t i) {
    const uint32_t *q = reinterpret_cast<const uint32_t *>(p + i);
    return EventDesc{sload(q, 0), sload(q, 1), sload(q, 2), sload(q, 3)};
}

// Workgroup-wide OR of two flags (any live 1 bit, any live 0 bit) through
// LDS: every lane of the workgroup calls it with its own flags and gets the
// workgroup's summary of the chunk.
__device__ __forceinline__ uint32_t chunk_summary(bool any1, bool any0) {
    __shared__ uint32_t flags[kLanes / 64];
    const uint32_t w = (__builtin_amdgcn_ballot_w64(any1) != 0 ? 1u : 0u) |
                       (__builtin_amdgcn_ballot_w64(any0) != 0 ? 2u : 0u);
    if ((threadIdx.x & 63u) == 0) flags[threadIdx.x >> 6] = w;
    __syncthreads();
    uint32_t f = 0;
#pragma unroll
    for (uint32_t i = 0; i < kLanes / 64; ++i) f |= flags[i];
    __syncthreads();  // flags[] is reused by the workgroup's next task
    // any 1 and any 0: mixed; only 1s: all ones; otherwise all zero.
    return f == 3u ? kChunkMixed : (f == 1u ? kChunkOnes : kChunkZero);
}

__device__ __forceinline__ void put_summary(const PassLaunch &L, uint32_t row, uint32_t chunk,
                                            uint32_t st) {
    if (threadIdx.x == 0) L.state[uint64_t(row) * L.state_stride + chunk] = uint8_t(st);
}

// Sampler task: words [base, base + chunk_words) of one event.
__device__ __forceinline__ void sampler_task(const PassLaunch &L, uint64_t *__restrict__ words,
                                             const LaunchGeometry &geo, const SplitFracs &sf,
                                             uint64_t task) {
    const uint32_t e = uint32_t(div_by_reciprocal(task, geo.n_chunks, sf.inv_chunks));
    const uint32_t chunk = uint32_t(task - uint64_t(e) * geo.n_chunks);
    const EventDesc d = event_desc(L.events, e);
    const uint32_t iteration =
        L.iteration_ptr ? sload(L.iteration_ptr, 0) + L.samp_iteration : L.samp_iteration;
    const uint32_t c3base = (iteration + 1u) << 6;
    uint64_t *row = words + uint64_t(L.samp_row0 + e) * geo.stride;
    const uint64_t base = uint64_t(chunk) * geo.chunk_words;
    const bool full = base + geo.chunk_words <= geo.n_bp;  // no padding word
    uint64_t any = 0, all = ~0ull;  // over the lane's live words
    for (uint32_t off = threadIdx.x * 2u; off < geo.chunk_words; off += 2u * kLanes) {
        const uint64_t s = base + off;
        uint64_t w0, w1;
        if (d.mode == kEventDraw) {
            w0 = bernoulli_word(d.ctr0, d.ctr1, uint32_t(s) + 1u, c3base, L.key0, L.key1, d.threshold);
            w1 = bernoulli_word(d.ctr0, d.ctr1, uint32_t(s) + 2u, c3base, L.key0, L.key1, d.threshold);
        } else {
            w0 = w1 = (d.mode == kEventAlways) ? ~0ull : 0ull;
        }
        store_pair(row + s, w0, w1);
        if (L.state) {
            if (full) {
                any |= w0 | w1;
                all &= w0 & w1;
            } else {
                if (s < geo.n_bp) { any |= w0; all &= w0; }
                if (s + 1 < geo.n_bp) { any |= w1; all &= w1; }
            }
        }
    }
    if (L.state) {
        const uint32_t st = chunk_summary(any != 0, all != ~0ull);
        put_summary(L, L.samp_row0 + e, chunk, st);
    }
}

struct Pair {
    uint64_t x, y;
};

// A gate task's view of the words matrix. Each row input is read through a
// buffer resource built on the scalar unit for the task's chunk of that row
// (row index read once into an SGPR), with the lane's byte offsets for the K
// pairs held in VGPRs for the whole task, so an input costs no vector
// address arithmetic. Lane l holds words base + 2 l + 512 k, k = 0..K-1, at
// byte offset 16 l + 4096 k of the chunk; the resource's size is the chunk's,
// so no access can leave it.
struct Task {
    const PassLaunch *L;
    uint64_t *__restrict__ words;
    unsigned long long *__restrict__ ones;
    uint64_t stride;      // words per row
    uint64_t base;        // first word of the chunk
    uint64_t n_bp;        // live words per row
    uint32_t lane_off;    // 16 * lane
    uint32_t chunk_bytes; // 8 * chunk words
    bool full;            // every word of the chunk is below n_bp (no padding)
    uint32_t chunk;       // chunk index (summary column)
    uint32_t live;        // live words of the chunk
};

constexpr int32_t kBufferWord3 = 0x00020000;  // raw buffer, 32-bit format fields (gfx9)
constexpr int32_t kNonTemporal = 2;           // cache policy: no temporal reuse expected

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t chunk_of(const Task &t, uint32_t row) {
    return __builtin_amdgcn_make_buffer_rsrc(t.words + (uint64_t(row) * t.stride + t.base), 0,
                                             int32_t(t.chunk_bytes), kBufferWord3);
}

__device__ __forceinline__ Pair load_k(__amdgpu_buffer_rsrc_t r, const Task &t, int k) {
    const u64x2 v = __builtin_bit_cast(
        u64x2, __builtin_amdgcn_raw_buffer_load_b128(r, int32_t(t.lane_off + 4096u * uint32_t(k)),
                                                     0, kNonTemporal));
    return {v.x, v.y};
}

// Whether word h of the lane's pair k is below n_bp: its offset in the
// chunk is 2 lane + 512 k + h, and the chunk's first `live` words are live.
__device__ __forceinline__ bool lane_word_live(const Task &t, int k, int h) {
    return 2u * threadIdx.x + 512u * uint32_t(k) + uint32_t(h) < t.live;
}

// One bit count per 32 bits, accumulated (v_bcnt_u32_b32 adds its second
// operand).
__device__ __forceinline__ uint32_t popc_add(uint64_t w, uint32_t acc) {
    return uint32_t(__popcll(w)) + acc;
}

// a & b & c, one bitop3 per 32 bits.
__device__ __forceinline__ uint64_t and3_64(uint64_t a, uint64_t b, uint64_t c) {
    const uint32_t lo = __builtin_amdgcn_bitop3_b32(uint32_t(a), uint32_t(b), uint32_t(c), 0x80);
    const uint32_t hi =
        __builtin_amdgcn_bitop3_b32(uint32_t(a >> 32), uint32_t(b >> 32), uint32_t(c >> 32), 0x80);
    return uint64_t(hi) << 32 | lo;
}

// Popcount of the lane's pairs k0 .. k0+K-1 and, with kSummary, the OR and
// AND of their live words (a chunk summary's inputs). A chunk without
// padding words (every chunk of a row whose words per pass are a multiple of
// the chunk, and every chunk but the last otherwise) takes the unmasked
// path; the branch is on a workgroup-uniform flag, so the common case tests
// no word for liveness.
template <int K, bool kSummary>
__device__ __forceinline__ uint32_t pair_stats(const Task &t, const Pair (&r)[K], int k0,
                                               uint64_t &any, uint64_t &all) {
    uint32_t pc = 0;
    if (t.full) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            pc = popc_add(r[k].y, popc_add(r[k].x, pc));
            if (kSummary) {
                any |= r[k].x | r[k].y;  // v_or3
                all = and3_64(all, r[k].x, r[k].y);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (lane_word_live(t, k0 + k, 0)) {
                pc = popc_add(r[k].x, pc);
                if (kSummary) { any |= r[k].x; all &= r[k].x; }
            }
            if (lane_word_live(t, k0 + k, 1)) {
                pc = popc_add(r[k].y, pc);
                if (kSummary) { any |= r[k].y; all &= r[k].y; }
            }
        }
    }
    return pc;
}

// Summary of the task's chunk of a row (kChunkMixed without summaries).
__device__ __forceinline__ uint32_t row_summary(const Task &t, uint32_t row) {
    if (!t.L->state) return kChunkMixed;
    const uint64_t b = uint64_t(row) * t.L->state_stride + t.chunk;
    const uint32_t w = sload(reinterpret_cast<const uint32_t *>(t.L->state), b >> 2);
    return (w >> (8u * uint32_t(b & 3u))) & 0xFFu;
}

// Classification of a gate over the task's chunk from its inputs'
// summaries: constant (every live word 0 or all ones, without loading an
// input), or the work left: the inputs still to fold (`mixed`), the XOR
// parity of the constant inputs, and ATLEAST's threshold over the mixed
// inputs (the constant true inputs already counted). An ATLEAST with 256 or
// more inputs keeps every input, since its 8-bit count wraps.
struct Cls {
    int32_t konst;   // -1: evaluate; 0 or 1: the gate's value (after NOT/NAND/NOR)
    uint32_t k8;     // ATLEAST threshold over the mixed inputs
    uint32_t flip;   // XOR: parity of the constant true inputs
    uint32_t mixed;  // inputs left to fold
};

// Value of one input over the chunk before the gate folds it: 0, 1 or
// kChunkMixed. With kFusedArgs an input may be a fused gate, whose value
// comes from its own inputs' summaries (fused gates never hold fused gates,
// so the classification has two levels).
template <bool kFusedArgs>
__device__ __forceinline__ uint32_t arg_value(uint32_t ai, uint32_t neg, const Task &t);

// 1 if input i is complemented (positive inputs come first, n_pos of them),
// else 0, in integer arithmetic: a comparison's boolean would be a lane
// mask, which costs a vector select and a read-back per input to turn into
// a number.
__device__ __forceinline__ uint32_t negated(uint32_t i, uint32_t n_pos) {
    return uint32_t(int32_t(n_pos - 1u - i) >> 31) & 1u;
}

template <bool kFusedArgs>
__device__ __forceinline__ Cls classify(uint32_t type, uint32_t k8, uint32_t n_in, uint32_t n_pos,
                                        const uint32_t *__restrict__ a, const Task &t) {
    Cls c{-1, k8, 0, n_in};
    if (!t.L->state || (type == 2u && n_in >= 256u)) return c;
    uint32_t n0 = 0, n1 = 0, last = kChunkZero;
    for (uint32_t i = 0; i < n_in; ++i) {
        last = arg_value<kFusedArgs>(sload(a, i), negated(i, n_pos), t);
        n0 += last == 0u;
        n1 += last == 1u;
    }
    const uint32_t m = n_in - n0 - n1;
    c.mixed = m;
    int32_t v = -1;
    if (type == 0u || type == 5u) {         // AND, NAND
        v = n0 ? 0 : (m == 0 ? 1 : -1);
    } else if (type == 1u || type == 6u) {  // OR, NOR
        v = n1 ? 1 : (m == 0 ? 0 : -1);
    } else if (type == 3u) {                // XOR
        c.flip = n1 & 1u;
        v = m == 0 ? int32_t(c.flip) : -1;
    } else if (type == 4u || type == 7u) {  // NOT, NULL: the last input
        v = n_in == 0 ? 0 : (last == kChunkMixed ? -1 : int32_t(last));
        c.mixed = v < 0 ? 1u : 0u;
    } else {                                // ATLEAST, fewer than 256 inputs
        const int32_t kk = int32_t(k8) - int32_t(n1);
        v = kk <= 0 ? 1 : (kk > int32_t(m) ? 0 : -1);
        c.k8 = kk > 0 ? uint32_t(kk) : 0u;
    }
    if (v >= 0 && (type == 4u || type == 5u || type == 6u)) v ^= 1;
    c.konst = v;
    return c;
}

template <bool kFusedArgs>
__device__ __forceinline__ uint32_t arg_value(uint32_t ai, uint32_t neg, const Task &t) {
    uint32_t v;
    if (kFusedArgs && (ai & kInlineArg)) {
        const GateDesc c = gate_desc(t.L->tree_gates, ai & ~kInlineArg);
        const Cls f = classify<false>(c.type_k_flags & 0xFFu, (c.type_k_flags >> 8) & 0xFFu,
                                      c.counts & 0xFFFFu, c.counts >> 16,
                                      t.L->arg_rows + c.arg_begin, t);
        v = f.konst < 0 ? kChunkMixed : uint32_t(f.konst);
    } else {
        v = row_summary(t, ai);
    }
    return v == kChunkMixed ? v : (v ^ neg);
}

// Whether input ai (a row or a fused gate) is folded: false when its chunk
// is constant (a constant the classification has already accounted for).
template <bool kFusedArgs>
__device__ __forceinline__ bool folded(uint32_t ai, const Task &t) {
    return !t.L->state || arg_value<kFusedArgs>(ai, false, t) == kChunkMixed;
}

// The tally of a constant chunk: live words x 64 one bits, or none; added
// once per workgroup.
__device__ __forceinline__ void tally_constant(const Task &t, uint32_t slot, int32_t v) {
    if (threadIdx.x == 0 && v == 1)
        atomicAdd(t.ones + slot, (unsigned long long)t.live * 64ull);
}

// Truth table of v_bitop3_b32 for acc <op> v0' <op> v1', v' = v or ~v; the
// table is indexed by 4 acc + 2 v0 + v1 (operand order S0, S1, S2). With
// kTwo false, v1 is ignored (a one-input fold).
template <int Op, bool kN0, bool kN1, bool kTwo>
constexpr uint32_t fold_table() {
    uint32_t t = 0;
    for (int i = 0; i < 8; ++i) {
        const bool acc = (i >> 2) & 1;
        const bool a = (((i >> 1) & 1) != 0) != kN0;
        const bool b = ((i & 1) != 0) != kN1;
        bool r;
        if (Op == 0) r = acc && a && (!kTwo || b);
        else if (Op == 1) r = acc || a || (kTwo && b);
        else r = acc != a != (kTwo && b);
        t |= uint32_t(r) << i;
    }
    return t;
}

template <uint32_t T>
__device__ __forceinline__ uint64_t bitop3_64(uint64_t a, uint64_t b, uint64_t c) {
    const uint32_t lo = __builtin_amdgcn_bitop3_b32(uint32_t(a), uint32_t(b), uint32_t(c), T);
    const uint32_t hi =
        __builtin_amdgcn_bitop3_b32(uint32_t(a >> 32), uint32_t(b >> 32), uint32_t(c >> 32), T);
    return uint64_t(hi) << 32 | lo;
}

// r <op>= v0' <op> v1': one bitop3 per 32 bits for two inputs.
template <int Op, bool kN0, bool kN1, int K>
__device__ __forceinline__ void fold2(Pair (&r)[K], const Pair (&v0)[K], const Pair (&v1)[K]) {
    constexpr uint32_t T = fold_table<Op, kN0, kN1, true>();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        r[k].x = bitop3_64<T>(r[k].x, v0[k].x, v1[k].x);
        r[k].y = bitop3_64<T>(r[k].y, v0[k].y, v1[k].y);
    }
}

template <int Op, int K>
__device__ __forceinline__ void fold1(Pair (&r)[K], const Pair (&v)[K], bool neg) {
    constexpr uint32_t TP = fold_table<Op, false, false, false>();
    constexpr uint32_t TN = fold_table<Op, true, false, false>();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (neg) {
            r[k].x = bitop3_64<TN>(r[k].x, v[k].x, v[k].x);
            r[k].y = bitop3_64<TN>(r[k].y, v[k].y, v[k].y);
        } else {
            r[k].x = bitop3_64<TP>(r[k].x, v[k].x, v[k].x);
            r[k].y = bitop3_64<TP>(r[k].y, v[k].y, v[k].y);
        }
    }
}

template <int K, int Op, bool kFused>
__device__ __forceinline__ void fold_simple(Pair (&r)[K], const uint32_t *__restrict__ a,
                                            uint32_t n_in, uint32_t n_pos, const Task &t,
                                            int k0 = 0, bool fused_args = false);

template <int K>
__device__ __forceinline__ void fill(Pair (&v)[K], uint64_t w) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = {w, w};
}

// Word pairs k0 .. k0+K-1 of a fused gate (AND/OR/XOR/NAND/NOR/NOT/NULL over
// rows), tallied: the popcount is added to the gate's counter here. With
// tally_constant_too false, a constant fused gate is left to the caller,
// which has counted it already.
template <int K>
__device__ __forceinline__ void eval_fused(uint32_t ref, const Task &t, Pair (&v)[K], int k0 = 0) {
    const GateDesc c = gate_desc(t.L->tree_gates, ref & ~kInlineArg);
    const uint32_t *ca = t.L->arg_rows + c.arg_begin;
    const uint32_t n_in = c.counts & 0xFFFFu, n_pos = c.counts >> 16;
    const uint32_t type = c.type_k_flags & 0xFFu;
    const Cls cls = classify<false>(type, 0u, n_in, n_pos, ca, t);
    if (cls.konst >= 0) {  // its tally was counted where the constant was found
        fill<K>(v, cls.konst ? ~0ull : 0ull);
        return;
    }
    if (type == 4u || type == 7u) {  // NOT, NULL: the last input
        fill<K>(v, 0ull);
        if (n_in) {
            const __amdgpu_buffer_rsrc_t rp = chunk_of(t, uniform(sload(ca, n_in - 1)));
#pragma unroll
            for (int k = 0; k < K; ++k) {
                v[k] = load_k(rp, t, k0 + k);
                if (n_in - 1 >= n_pos) { v[k].x = ~v[k].x; v[k].y = ~v[k].y; }
            }
        }
    } else if (type == 0u || type == 5u) {
        fill<K>(v, ~0ull);
        fold_simple<K, 0, false>(v, ca, n_in, n_pos, t, k0);
    } else if (type == 3u) {
        fill<K>(v, cls.flip ? ~0ull : 0ull);
        fold_simple<K, 2, false>(v, ca, n_in, n_pos, t, k0);
    } else {
        fill<K>(v, 0ull);
        fold_simple<K, 1, false>(v, ca, n_in, n_pos, t, k0);
    }
    const bool invert = type == 4u || type == 5u || type == 6u;  // NOT, NAND, NOR
    if (invert) {
#pragma unroll
        for (int k = 0; k < K; ++k) { v[k].x = ~v[k].x; v[k].y = ~v[k].y; }
    }
    if (c.type_k_flags & kGateTallied) {
        uint64_t any = 0, all = ~0ull;
        tally_wave(t.ones, c.out_row, pair_stats<K, false>(t, v, k0, any, all));
    }
}

// The tallies of the fused gates among a task's inputs whose chunks are
// constant: their words are never computed, so their counts are added here,
// once per workgroup. Called once per task.
__device__ __forceinline__ void tally_constant_fused(const uint32_t *__restrict__ a, uint32_t n_in,
                                                     const Task &t) {
    if (!t.L->state) return;
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint32_t ai = sload(a, i);
        if (!(ai & kInlineArg)) continue;
        const GateDesc c = gate_desc(t.L->tree_gates, ai & ~kInlineArg);
        if (!(c.type_k_flags & kGateTallied)) continue;
        const uint32_t v = arg_value<true>(ai, false, t);
        if (v != kChunkMixed) tally_constant(t, c.out_row, int32_t(v));
    }
}

// AND / OR / XOR families over K pairs per lane: the inputs [i0, i1) of one
// sign (kNeg: complemented). Inputs whose chunks are constant are skipped:
// the classification found none that decides the gate, so each is neutral
// (AND: true, OR: false) or, for XOR, part of the parity the caller starts
// from. With kFused an argument may reference a fused gate (kInlineArg),
// whose word the task computes in registers; fused gates never reference
// fused gates. The folds are associative and commutative bit by bit, so the
// order in which inputs are folded changes no word.
template <int K, int Op, bool kNeg, bool kFused>
__device__ __forceinline__ void fold_range(Pair (&r)[K], const uint32_t *__restrict__ a,
                                           uint32_t i0, uint32_t i1, const Task &t, int k0,
                                           bool fused_args) {
    if (!kFused) {
        // Inside a fused gate, whose consumer's accumulator is live: one row
        // at a time (register pressure).
        for (uint32_t i = i0; i < i1; ++i) {
            const uint32_t ai = uniform(sload(a, i));
            if (!folded<false>(ai, t)) continue;
            const __amdgpu_buffer_rsrc_t p0 = chunk_of(t, ai);
            Pair v[K];
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] = load_k(p0, t, k0 + k);
            fold1<Op, K>(r, v, kNeg);
        }
        return;
    }
    // Rows two at a time (2K 16-byte loads in flight, one bitop3 per 32 bits
    // for both): a row waits in `held` (a scalar) for the next one, so the
    // loop body has this one path and the accumulators stay in their
    // registers from one iteration to the next.
    uint32_t held = kInlineArg;  // none
    for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t ai = uniform(sload(a, i));
        if ((ai & kInlineArg) || !folded<false>(ai, t)) continue;
        if (held == kInlineArg) {
            held = ai;
            continue;
        }
        const __amdgpu_buffer_rsrc_t p0 = chunk_of(t, held), p1 = chunk_of(t, ai);
        Pair v0[K], v1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            v0[k] = load_k(p0, t, k0 + k);
            v1[k] = load_k(p1, t, k0 + k);
        }
        fold2<Op, kNeg, kNeg>(r, v0, v1);
        held = kInlineArg;
    }
    if (held != kInlineArg) {
        const __amdgpu_buffer_rsrc_t p0 = chunk_of(t, held);
        Pair v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = load_k(p0, t, k0 + k);
        fold1<Op, K>(r, v, kNeg);
    }
    // Fused inputs (only when the gate has some: kGateHasFused): each one's
    // word computed here, then folded in.
    if (!fused_args) return;
    for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t ai = uniform(sload(a, i));
        if (!(ai & kInlineArg) || !folded<true>(ai, t)) continue;
        Pair v[K];
        eval_fused<K>(ai, t, v);
        fold1<Op, K>(r, v, kNeg);
    }
}

// Positive inputs come first (n_pos of them), then the complemented ones.
template <int K, int Op, bool kFused>
__device__ __forceinline__ void fold_simple(Pair (&r)[K], const uint32_t *__restrict__ a,
                                            uint32_t n_in, uint32_t n_pos, const Task &t,
                                            int k0, bool fused_args) {
    fold_range<K, Op, false, kFused>(r, a, 0, n_pos, t, k0, fused_args);
    fold_range<K, Op, true, kFused>(r, a, n_pos, n_in, t, k0, fused_args);
}

// Bit-sliced counters: plane b of a counter holds bit b of every trial's
// count. Adding two input words takes one full adder at plane 0 (sum and
// majority, one bitop3 each per 32 bits) and a ripple of its carry through
// the planes above (two operations per plane, one at the top, whose carry
// out is dropped: the count of at most 2^P - 1 inputs never needs it, and
// an 8-plane counter wraps as the reference's uint8 count does). Adding
// inputs one at a time took two operations per plane for each.
template <int P>
__device__ __forceinline__ void count_add2(uint64_t (&c)[P], uint64_t a, uint64_t b) {
    uint64_t carry = bitop3_64<0xE8>(c[0], a, b);  // majority(c0, a, b)
    c[0] = bitop3_64<0x96>(c[0], a, b);            // c0 ^ a ^ b
#pragma unroll
    for (int p = 1; p + 1 < P; ++p) {
        const uint64_t t = c[p] & carry;
        c[p] ^= carry;
        carry = t;
    }
    c[P - 1] ^= carry;
}

template <int P>
__device__ __forceinline__ void count_add1(uint64_t (&c)[P], uint64_t a) {
    uint64_t carry = a;
#pragma unroll
    for (int p = 0; p + 1 < P; ++p) {
        const uint64_t t = c[p] & carry;
        c[p] ^= carry;
        carry = t;
    }
    c[P - 1] ^= carry;
}

// At most 15 inputs folded, pair k: four planes hold the count exactly; a
// threshold above 15 can never be met (the comparison sees a k bit above
// the planes as unmet). Inputs with constant chunks are skipped (k8 already
// counts the constant true ones). Inputs are added two at a time: one is
// held until the next arrives.
template <bool kFused>
__device__ __forceinline__ Pair atleast_pair4(const uint32_t *__restrict__ a, uint32_t n_in,
                                              uint32_t n_pos, uint32_t k8, const Task &t, int k) {
    uint64_t cx[4] = {0, 0, 0, 0}, cy[4] = {0, 0, 0, 0};
    Pair held{0, 0};
    bool holding = false;  // wave-uniform
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint32_t ai = uniform(sload(a, i));
        if (!folded<kFused>(ai, t)) continue;
        Pair v;
        if (kFused && (ai & kInlineArg)) {
            Pair f[1];
            eval_fused<1>(ai, t, f, k);  // pair k only; tallied pair by pair
            v = f[0];
        } else {
            v = load_k(chunk_of(t, ai), t, k);
        }
        if (i >= n_pos) { v.x = ~v.x; v.y = ~v.y; }
        if (!holding) {
            held = v;
            holding = true;
            continue;
        }
        count_add2<4>(cx, held.x, v.x);
        count_add2<4>(cy, held.y, v.y);
        holding = false;
    }
    if (holding) {
        count_add1<4>(cx, held.x);
        count_add1<4>(cy, held.y);
    }
    if (k8 > 15u) return {0, 0};
    return {atleast_compare<uint64_t, 4>(cx, k8), atleast_compare<uint64_t, 4>(cy, k8)};
}

// Any fan-in (row inputs only), word h of pair k: eight planes hold the
// reference's 8-bit count, which wraps at 256 as its uint8 counter does.
// Below 256 inputs constant chunks are skipped (k8 counts the true ones);
// from 256 on every input is counted, a constant chunk as its constant.
__device__ __forceinline__ uint64_t atleast_word8(const uint32_t *__restrict__ a, uint32_t n_in,
                                                  uint32_t n_pos, uint32_t k8, const Task &t, int k,
                                                  int h) {
    uint64_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool wide = n_in >= 256u;
    uint64_t held = 0;
    bool holding = false;  // wave-uniform
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint32_t ai = uniform(sload(a, i));
        const uint32_t st = row_summary(t, ai);
        if (st != kChunkMixed && !wide) continue;
        uint64_t v = st == kChunkZero ? 0ull : ~0ull;
        if (st == kChunkMixed)
            v = __builtin_bit_cast(
                uint64_t, __builtin_amdgcn_raw_buffer_load_b64(
                              chunk_of(t, ai),
                              int32_t(t.lane_off + 4096u * uint32_t(k) + 8u * uint32_t(h)), 0,
                              kNonTemporal));
        if (i >= n_pos) v = ~v;
        if (!holding) {
            held = v;
            holding = true;
            continue;
        }
        count_add2<8>(c, held, v);
        holding = false;
    }
    if (holding) count_add1<8>(c, held);
    return atleast_compare<uint64_t, 8>(c, k8);
}

// The gate's words for pairs k = 0..K-1 (no store, no tally of its own),
// for a gate that classification left to evaluate. Fused gates appear only
// under AND/OR/XOR/NAND/NOR and ATLEAST with at most 15 inputs, never under
// NOT/NULL or a wide ATLEAST.
template <int K, bool kFused>
__device__ __forceinline__ void eval_gate(const GateDesc &d, const uint32_t *__restrict__ a,
                                          const Task &t, const Cls &cls, Pair (&r)[K]) {
    const uint32_t n_in = d.counts & 0xFFFFu;
    const uint32_t n_pos = d.counts >> 16;
    const uint32_t type = d.type_k_flags & 0xFFu;
    const bool fused_args = kFused && (d.type_k_flags & kGateHasFused);
    if (type == 2u) {  // ATLEAST
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (n_in <= 15u || (n_in < 256u && cls.mixed <= 15u))
                r[k] = atleast_pair4<kFused>(a, n_in, n_pos, cls.k8, t, k);
            else
                r[k] = {atleast_word8(a, n_in, n_pos, cls.k8, t, k, 0),
                        atleast_word8(a, n_in, n_pos, cls.k8, t, k, 1)};
        }
    } else if (type == 4u || type == 7u) {  // NOT, NULL: the last input
        fill<K>(r, 0ull);
        if (n_in) {
            const __amdgpu_buffer_rsrc_t rp = chunk_of(t, uniform(sload(a, n_in - 1)));
#pragma unroll
            for (int k = 0; k < K; ++k) {
                r[k] = load_k(rp, t, k);
                if (n_in - 1 >= n_pos) { r[k].x = ~r[k].x; r[k].y = ~r[k].y; }
            }
        }
    } else {
        const bool and_like = (type == 0u || type == 5u);
        fill<K>(r, and_like ? ~0ull : (cls.flip ? ~0ull : 0ull));
        if (and_like)
            fold_simple<K, 0, kFused>(r, a, n_in, n_pos, t, 0, fused_args);
        else if (type == 3u)
            fold_simple<K, 2, kFused>(r, a, n_in, n_pos, t, 0, fused_args);
        else
            fold_simple<K, 1, kFused>(r, a, n_in, n_pos, t, 0, fused_args);
    }
    if (type == 4u || type == 5u || type == 6u) {  // NOT, NAND, NOR
#pragma unroll
        for (int k = 0; k < K; ++k) { r[k].x = ~r[k].x; r[k].y = ~r[k].y; }
    }
}

template <int K>
__device__ __forceinline__ void store_k(const Task &t, uint32_t row, const Pair (&r)[K]) {
    const __amdgpu_buffer_rsrc_t o = chunk_of(t, row);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        u64x2 v;
        v.x = r[k].x;
        v.y = r[k].y;
        __builtin_amdgcn_raw_buffer_store_b128(
            __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), o,
            int32_t(t.lane_off + 4096u * uint32_t(k)), 0, kNonTemporal);
    }
}

template <int K>
__device__ __forceinline__ void gate_task_k(const GateDesc &d, const uint32_t *__restrict__ a,
                                            const Task &t) {
    const uint32_t n_in = d.counts & 0xFFFFu;
    const uint32_t out = uniform(d.out_row);
    const bool store = !(d.type_k_flags & kGateNoStore);
    const bool summarise = t.L->state && store;
    const Cls cls = classify<true>(d.type_k_flags & 0xFFu, (d.type_k_flags >> 8) & 0xFFu, n_in,
                                   d.counts >> 16, a, t);
    const bool fused_args = d.type_k_flags & kGateHasFused;
    if (fused_args) tally_constant_fused(a, n_in, t);
    if (cls.konst >= 0) {
        // Decided by summaries alone: no input is loaded. Fused inputs whose
        // chunks are not constant are still evaluated, for their tallies.
        for (uint32_t i = 0; fused_args && i < n_in; ++i) {
            const uint32_t ai = uniform(sload(a, i));
            if ((ai & kInlineArg) && folded<true>(ai, t)) {
                Pair v[K];
                eval_fused<K>(ai, t, v);
            }
        }
        if (d.type_k_flags & kGateTallied) tally_constant(t, out, cls.konst);
        if (d.type_k_flags & kGateForceStore) {
            Pair r[K];
            fill<K>(r, cls.konst ? ~0ull : 0ull);
            store_k<K>(t, out, r);
        }
        if (summarise) put_summary(*t.L, out, t.chunk, uint32_t(cls.konst));
        return;
    }
    Pair r[K];
    eval_gate<K, true>(d, a, t, cls, r);
    uint64_t any = 0, all = ~0ull;
    const uint32_t pc = summarise ? pair_stats<K, true>(t, r, 0, any, all)
                                  : pair_stats<K, false>(t, r, 0, any, all);
    if (d.type_k_flags & kGateTallied) tally_wave(t.ones, out, pc);
    if (!store) return;
    uint32_t st = kChunkMixed;
    if (summarise) {
        st = chunk_summary(any != 0, all != ~0ull);
        put_summary(*t.L, out, t.chunk, st);
    }
    if (st == kChunkMixed || (d.type_k_flags & kGateForceStore)) store_k<K>(t, out, r);
}

// Gate task: one gate over words [base, base + chunk_words).
__device__ __forceinline__ void gate_task(const PassLaunch &L, uint64_t *__restrict__ words,
                                          unsigned long long *__restrict__ ones,
                                          const LaunchGeometry &geo, const SplitFracs &sf,
                                          uint64_t task) {
    const uint32_t chunk = uint32_t(div_by_reciprocal(task, L.n_gates, sf.inv_gates));
    const uint32_t gi = uint32_t(task - uint64_t(chunk) * L.n_gates);
    const GateDesc d = gate_desc(L.gates, gi);
    const uint32_t *a = L.arg_rows + d.arg_begin;
    const uint64_t base = uint64_t(chunk) * geo.chunk_words;
    const uint64_t live = geo.n_bp - base < geo.chunk_words ? geo.n_bp - base : geo.chunk_words;
    const Task t{&L, words, ones, geo.stride, base, geo.n_bp, 16u * threadIdx.x,
                 8u * geo.chunk_words, base + geo.chunk_words <= geo.n_bp, chunk,
                 uint32_t(live)};
    switch (geo.chunk_words) {
    case 2048: gate_task_k<4>(d, a, t); break;
    case 1024: gate_task_k<2>(d, a, t); break;
    default: gate_task_k<1>(d, a, t); break;
    }
}

}  // namespace

__global__ __launch_bounds__(256) void pass_layer(PassLaunch L, uint64_t *__restrict__ words,
                                                  unsigned long long *__restrict__ ones,
                                                  LaunchGeometry geo, SplitFracs sf) {
    const uint64_t ng = gate_tasks(L, geo.n_chunks);
    const uint64_t ns = L.samp_end - L.samp_begin;
    const uint32_t x = blockIdx.x % kXcds;
    const uint64_t j = blockIdx.x / kXcds;
    const XcdShare sh = xcd_share(ng, ns, x);
    if (j >= sh.tx) return;
    const uint64_t before = gates_before(j, sh.ngx, sf.frac[x]);
    const bool is_gate = gates_before(j + 1, sh.ngx, sf.frac[x]) > before;
    uint64_t task = is_gate ? sh.g_lo + before : L.samp_begin + sh.s_lo + (j - before);
    // The task's parameter sample: its words, summaries, counters, event
    // descriptors and key (sample 0, and nothing moves, in a one-sample launch).
    uint64_t s = 0;
    if (is_gate && L.n_samples > 1) {
        const uint64_t per = uint64_t(L.n_gates) * geo.n_chunks;
        s = div_by_reciprocal(task, per, sf.inv_sample_gate_tasks);
        task -= s * per;
    } else if (!is_gate && L.samp_events) {
        const uint64_t per = uint64_t(L.samp_events) * geo.n_chunks;
        s = div_by_reciprocal(task, per, sf.inv_sample_samp_tasks);
        task -= s * per;
    }
    PassLaunch Ls = L;
    if (s) {
        if (Ls.state) Ls.state += s * L.sample_state;
        Ls.events += s * L.samp_events;
        words += s * L.sample_words;
        ones += s * L.sample_ones;
    }
    if (!is_gate && L.keys0) Ls.key0 = sload(L.keys0, s);
    if (is_gate)
        gate_task(Ls, words, ones, geo, sf, task);
    else
        sampler_task(Ls, words, geo, sf, task);
}

uint64_t pass_grid_blocks(const PassLaunch &l, const LaunchGeometry &geo) {
    const uint64_t ng = gate_tasks(l, geo.n_chunks);
    const uint64_t ns = l.samp_end - l.samp_begin;
    uint64_t m = 0;
    for (uint32_t x = 0; x < kXcds; ++x) {
        const uint64_t t = xcd_share(ng, ns, x).tx;
        m = t > m ? t : m;
    }
    return m * kXcds;
}

hipError_t launch_pass(hipStream_t stream, const PassLaunch &l, uint64_t *words,
                       unsigned long long *ones, const LaunchGeometry &geo) {
    const uint64_t blocks = pass_grid_blocks(l, geo);
    if (blocks == 0) return hipSuccess;
    if (blocks >= (uint64_t(1) << 24)) return hipErrorInvalidConfiguration;
    const uint64_t ng = gate_tasks(l, geo.n_chunks);
    const uint64_t ns = l.samp_end - l.samp_begin;
    SplitFracs sf;
    for (uint32_t x = 0; x < kXcds; ++x) {
        const XcdShare sh = xcd_share(ng, ns, x);
        sf.frac[x] = interleave_frac(sh.ngx, sh.tx);
    }
    sf.inv_gates = reciprocal64(l.n_gates);
    sf.inv_chunks = reciprocal64(geo.n_chunks);
    sf.inv_sample_gate_tasks = reciprocal64(uint64_t(l.n_gates) * geo.n_chunks);
    sf.inv_sample_samp_tasks = reciprocal64(uint64_t(l.samp_events) * geo.n_chunks);
    hipLaunchKernelGGL(pass_layer, dim3(uint32_t(blocks)), dim3(kLanes), 0, stream, l, words, ones,
                       geo, sf);
    return hipGetLastError();
}

}  // namespace mcs
