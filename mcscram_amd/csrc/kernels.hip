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
constexpr uint32_t kXcds = 8;  // workgroups are dealt to the 8 XCDs round-robin

__device__ __forceinline__ uint32_t live_popcount(uint64_t w0, uint64_t w1, uint64_t s,
                                                  uint64_t n_bp) {
    return (s < n_bp ? __popcll(w0) : 0u) + (s + 1 < n_bp ? __popcll(w1) : 0u);
}

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

__device__ __forceinline__ void tally_wave(unsigned long long *ones, uint32_t slot, uint32_t pc) {
    pc = wave_sum(pc);
    if ((threadIdx.x & 63u) == 0 && pc) atomicAdd(ones + slot, (unsigned long long)pc);
}

// Sampler task: words [base, base + chunk_words) of one event.
__device__ __forceinline__ void sampler_task(const PassLaunch &L, uint64_t *__restrict__ words,
                                             const LaunchGeometry &geo, uint64_t task) {
    const uint32_t e = uint32_t(task / geo.n_chunks);
    const uint32_t chunk = uint32_t(task - uint64_t(e) * geo.n_chunks);
    const EventDesc d = L.events[e];
    const uint32_t iteration =
        L.iteration_ptr ? *L.iteration_ptr + L.samp_iteration : L.samp_iteration;
    const uint32_t c3base = (iteration + 1u) << 6;
    uint64_t *row = words + uint64_t(L.samp_row0 + e) * geo.stride;
    const uint64_t base = uint64_t(chunk) * geo.chunk_words;
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
    }
}

struct Pair {
    uint64_t x, y;
};

__device__ __forceinline__ Pair load_pair(const uint64_t *__restrict__ p) {
    const u64x2 v = __builtin_nontemporal_load(reinterpret_cast<const u64x2 *>(p));
    return {v.x, v.y};
}

// Inputs of a task. A row input is a 16-byte load per pair; with kFused an
// argument may instead reference a fused gate (kInlineArg), whose word the
// task computes in registers from that gate's own row inputs and counts in
// that gate's tally. Fused gates never reference fused gates.
struct Inputs {
    const PassLaunch *L;
    const uint64_t *__restrict__ words;
    unsigned long long *__restrict__ ones;
    const LaunchGeometry *geo;
};

template <int K, bool kFused>
__device__ __forceinline__ void eval_gate(const GateDesc &d, const uint32_t *__restrict__ a,
                                          const Inputs &in, uint64_t s0, Pair (&r)[K]);

template <int K, int Op, bool kFused>
__device__ __forceinline__ void fold_simple(Pair (&r)[K], const uint32_t *__restrict__ a,
                                            uint32_t n_in, uint32_t n_pos, const Inputs &in,
                                            uint64_t s0);

// Word pairs k = 0..K-1 (lane words s0 + 512 k) of a fused gate, tallied.
// Fused gates are AND/OR/XOR/NAND/NOR/NOT/NULL over rows.
template <int K>
__device__ __forceinline__ void eval_fused(uint32_t ref, const Inputs &in, uint64_t s0,
                                           Pair (&v)[K]) {
    const GateDesc c = in.L->tree_gates[ref & ~kInlineArg];
    const uint32_t *ca = in.L->arg_rows + c.arg_begin;
    const uint32_t n_in = c.counts & 0xFFFFu, n_pos = c.counts >> 16;
    const uint32_t type = c.type_k_flags & 0xFFu;
    if (type == 4u || type == 7u) {  // NOT, NULL: the last input
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = {0, 0};
        if (n_in) {
            const uint64_t *p = in.words + uint64_t(ca[n_in - 1]) * in.geo->stride + s0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                v[k] = load_pair(p + 2u * kLanes * k);
                if (n_in - 1 >= n_pos) { v[k].x = ~v[k].x; v[k].y = ~v[k].y; }
            }
        }
    } else if (type == 0u || type == 5u) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = {~0ull, ~0ull};
        fold_simple<K, 0, false>(v, ca, n_in, n_pos, in, s0);
    } else if (type == 3u) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = {0, 0};
        fold_simple<K, 2, false>(v, ca, n_in, n_pos, in, s0);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = {0, 0};
        fold_simple<K, 1, false>(v, ca, n_in, n_pos, in, s0);
    }
    const bool invert = type == 4u || type == 5u || type == 6u;  // NOT, NAND, NOR
    uint32_t pc = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (invert) { v[k].x = ~v[k].x; v[k].y = ~v[k].y; }
        pc += live_popcount(v[k].x, v[k].y, s0 + 2u * kLanes * k, in.geo->n_bp);
    }
    if (c.type_k_flags & kGateTallied) tally_wave(in.ones, c.out_row, pc);
}

// AND / OR / XOR families over K pairs per lane (lane words s0 + 512 k):
// row inputs two at a time so 2K 16-byte loads are in flight per lane.
template <int K, int Op, bool kFused>  // Op: 0 AND, 1 OR, 2 XOR
__device__ __forceinline__ void fold_simple(Pair (&r)[K], const uint32_t *__restrict__ a,
                                            uint32_t n_in, uint32_t n_pos, const Inputs &in,
                                            uint64_t s0) {
    auto apply = [&](Pair &acc, Pair v) {
        if (Op == 0) { acc.x &= v.x; acc.y &= v.y; }
        if (Op == 1) { acc.x |= v.x; acc.y |= v.y; }
        if (Op == 2) { acc.x ^= v.x; acc.y ^= v.y; }
    };
    const uint64_t stride = in.geo->stride;
    uint32_t i = 0;
    while (i < n_in) {
        if (kFused && (a[i] & kInlineArg)) {
            Pair v[K];
            eval_fused<K>(a[i], in, s0, v);
            const bool n0 = i >= n_pos;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (n0) { v[k].x = ~v[k].x; v[k].y = ~v[k].y; }
                apply(r[k], v[k]);
            }
            i += 1;
            continue;
        }
        if (i + 1 < n_in && !(kFused && (a[i + 1] & kInlineArg))) {
            const uint64_t *p0 = in.words + uint64_t(a[i]) * stride + s0;
            const uint64_t *p1 = in.words + uint64_t(a[i + 1]) * stride + s0;
            Pair v0[K], v1[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                v0[k] = load_pair(p0 + 2u * kLanes * k);
                v1[k] = load_pair(p1 + 2u * kLanes * k);
            }
            const bool n0 = i >= n_pos, n1 = i + 1 >= n_pos;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (n0) { v0[k].x = ~v0[k].x; v0[k].y = ~v0[k].y; }
                if (n1) { v1[k].x = ~v1[k].x; v1[k].y = ~v1[k].y; }
                apply(r[k], v0[k]);
                apply(r[k], v1[k]);
            }
            i += 2;
            continue;
        }
        const uint64_t *p0 = in.words + uint64_t(a[i]) * stride + s0;
        const bool n0 = i >= n_pos;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            Pair v = load_pair(p0 + 2u * kLanes * k);
            if (n0) { v.x = ~v.x; v.y = ~v.y; }
            apply(r[k], v);
        }
        i += 1;
    }
}

// Up to 15 inputs: four planes hold the count exactly; k8 above 15 can never
// be met (the comparison sees a k bit above the planes as unmet).
template <bool kFused>
__device__ __forceinline__ Pair atleast_pair4(const uint32_t *__restrict__ a, uint32_t n_in,
                                              uint32_t n_pos, uint32_t k8, const Inputs &in,
                                              uint64_t s) {
    uint64_t cx[4] = {0, 0, 0, 0}, cy[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < n_in; ++i) {
        Pair v;
        if (kFused && (a[i] & kInlineArg)) {
            Pair f[1];
            eval_fused<1>(a[i], in, s, f);
            v = f[0];
        } else {
            v = load_pair(in.words + uint64_t(a[i]) * in.geo->stride + s);
        }
        if (i >= n_pos) { v.x = ~v.x; v.y = ~v.y; }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t tx = cx[b] & v.x, ty = cy[b] & v.y;
            cx[b] ^= v.x;
            cy[b] ^= v.y;
            v.x = tx;
            v.y = ty;
        }
    }
    if (k8 > 15u) return {0, 0};
    return {atleast_compare<uint64_t, 4>(cx, k8), atleast_compare<uint64_t, 4>(cy, k8)};
}

// Any fan-in (row inputs only): eight planes hold the reference's 8-bit
// count, which wraps at 256 as its uint8 counter does; one word at a time to
// bound registers.
__device__ __forceinline__ uint64_t atleast_word8(const uint32_t *__restrict__ a, uint32_t n_in,
                                                  uint32_t n_pos, uint32_t k8, const Inputs &in,
                                                  uint64_t s) {
    uint64_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < n_in; ++i) {
        uint64_t v = in.words[uint64_t(a[i]) * in.geo->stride + s];
        if (i >= n_pos) v = ~v;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t t = c[b] & v;
            c[b] ^= v;
            v = t;
        }
    }
    return atleast_compare<uint64_t, 8>(c, k8);
}

// The gate's words for pairs k = 0..K-1 (no store, no tally of its own).
// Fused gates appear only under AND/OR/XOR/NAND/NOR and ATLEAST with at most
// 15 inputs, never under NOT/NULL or a wide ATLEAST.
template <int K, bool kFused>
__device__ __forceinline__ void eval_gate(const GateDesc &d, const uint32_t *__restrict__ a,
                                          const Inputs &in, uint64_t s0, Pair (&r)[K]) {
    const uint32_t n_in = d.counts & 0xFFFFu;
    const uint32_t n_pos = d.counts >> 16;
    const uint32_t type = d.type_k_flags & 0xFFu;
    const uint32_t k8 = (d.type_k_flags >> 8) & 0xFFu;
    if (type == 2u) {  // ATLEAST
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint64_t s = s0 + 2u * kLanes * k;
            if (n_in <= 15u)
                r[k] = atleast_pair4<kFused>(a, n_in, n_pos, k8, in, s);
            else
                r[k] = {atleast_word8(a, n_in, n_pos, k8, in, s),
                        atleast_word8(a, n_in, n_pos, k8, in, s + 1)};
        }
    } else if (type == 4u || type == 7u) {  // NOT, NULL: the last input
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = {0, 0};
        if (n_in) {
            const uint64_t *p = in.words + uint64_t(a[n_in - 1]) * in.geo->stride + s0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                r[k] = load_pair(p + 2u * kLanes * k);
                if (n_in - 1 >= n_pos) { r[k].x = ~r[k].x; r[k].y = ~r[k].y; }
            }
        }
    } else {
        const bool and_like = (type == 0u || type == 5u);
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = and_like ? Pair{~0ull, ~0ull} : Pair{0, 0};
        if (and_like)
            fold_simple<K, 0, kFused>(r, a, n_in, n_pos, in, s0);
        else if (type == 3u)
            fold_simple<K, 2, kFused>(r, a, n_in, n_pos, in, s0);
        else
            fold_simple<K, 1, kFused>(r, a, n_in, n_pos, in, s0);
    }
    if (type == 4u || type == 5u || type == 6u) {  // NOT, NAND, NOR
#pragma unroll
        for (int k = 0; k < K; ++k) { r[k].x = ~r[k].x; r[k].y = ~r[k].y; }
    }
}

template <int K>
__device__ __forceinline__ void gate_task_k(const GateDesc &d, const uint32_t *__restrict__ a,
                                            const Inputs &in, uint64_t *__restrict__ words,
                                            uint64_t base) {
    const uint64_t s0 = base + threadIdx.x * 2u;
    Pair r[K];
    eval_gate<K, true>(d, a, in, s0, r);
    if (!(d.type_k_flags & kGateNoStore)) {
        uint64_t *out = words + uint64_t(d.out_row) * in.geo->stride + s0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            store_pair(out + 2u * kLanes * k, r[k].x, r[k].y);
    }
    if (d.type_k_flags & kGateTallied) {
        uint32_t pc = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            pc += live_popcount(r[k].x, r[k].y, s0 + 2u * kLanes * k, in.geo->n_bp);
        tally_wave(in.ones, d.out_row, pc);
    }
}

// Gate task: one gate over words [base, base + chunk_words).
__device__ __forceinline__ void gate_task(const PassLaunch &L, uint64_t *__restrict__ words,
                                          unsigned long long *__restrict__ ones,
                                          const LaunchGeometry &geo, uint64_t task) {
    const uint32_t chunk = uint32_t(task / L.n_gates);
    const uint32_t gi = uint32_t(task - uint64_t(chunk) * L.n_gates);
    const GateDesc d = L.gates[gi];
    const uint32_t *a = L.arg_rows + d.arg_begin;
    const uint64_t base = uint64_t(chunk) * geo.chunk_words;
    const Inputs in{&L, words, ones, &geo};
    switch (geo.chunk_words) {
    case 2048: gate_task_k<4>(d, a, in, words, base); break;
    case 1024: gate_task_k<2>(d, a, in, words, base); break;
    default: gate_task_k<1>(d, a, in, words, base); break;
    }
}

// Per-XCD split of a launch: workgroup b runs on XCD b % 8 as its (b / 8)-th
// workgroup there. XCD x owns the x-th eighth of the gate tasks (chunk-major,
// so the gates it evaluates together read the same chunk of their input rows
// through its own L2) and the x-th eighth of the sampler tasks, and deals
// them out interleaved in proportion.
struct Split {
    uint64_t g_lo, s_lo, ngx, tx;
};

__host__ __device__ __forceinline__ Split split_for(uint64_t ng, uint64_t ns, uint32_t x) {
    Split s;
    s.g_lo = ng * x / kXcds;
    s.s_lo = ns * x / kXcds;
    s.ngx = ng * (x + 1) / kXcds - s.g_lo;
    s.tx = s.ngx + (ns * (x + 1) / kXcds - s.s_lo);
    return s;
}

}  // namespace

__global__ __launch_bounds__(256) void pass_layer(PassLaunch L, uint64_t *__restrict__ words,
                                                  unsigned long long *__restrict__ ones,
                                                  LaunchGeometry geo) {
    const uint64_t ng = uint64_t(L.n_gates) * geo.n_chunks;
    const uint64_t ns = L.samp_end - L.samp_begin;
    const uint32_t x = blockIdx.x % kXcds;
    const uint64_t j = blockIdx.x / kXcds;
    const Split sp = split_for(ng, ns, x);
    if (j >= sp.tx) return;
    const uint64_t before = j * sp.ngx / sp.tx;
    const uint64_t upto = (j + 1) * sp.ngx / sp.tx;
    if (upto > before)
        gate_task(L, words, ones, geo, sp.g_lo + before);
    else
        sampler_task(L, words, geo, L.samp_begin + sp.s_lo + (j - before));
}

uint64_t pass_grid_blocks(const PassLaunch &l, const LaunchGeometry &geo) {
    const uint64_t ng = uint64_t(l.n_gates) * geo.n_chunks;
    const uint64_t ns = l.samp_end - l.samp_begin;
    uint64_t m = 0;
    for (uint32_t x = 0; x < kXcds; ++x) {
        const uint64_t t = split_for(ng, ns, x).tx;
        m = t > m ? t : m;
    }
    return m * kXcds;
}

hipError_t launch_pass(hipStream_t stream, const PassLaunch &l, uint64_t *words,
                       unsigned long long *ones, const LaunchGeometry &geo) {
    const uint64_t blocks = pass_grid_blocks(l, geo);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(pass_layer, dim3(uint32_t(blocks)), dim3(kLanes), 0, stream, l, words, ones,
                       geo);
    return hipGetLastError();
}

}  // namespace mcs
