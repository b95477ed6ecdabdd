// HIP kernels of the Monte Carlo fault-tree engine, written for CDNA4 (gfx950).
//
//   sample_events  Philox4x32-10 Bernoulli sampler, one uint64 word = 64
//                  trials per (event, word); lanes own consecutive words of
//                  one event so every per-event value sits in SGPRs and the
//                  stores are coalesced 16-byte writes.
//   eval_gates     one launch per layer: every gate of the layer over a chunk
//                  of words, 16-byte loads of each input row, bit-sliced
//                  counter for ATLEAST, popcount tally fused at production.
//
// Reference semantics restated here:
//   counter (pdag_idx+1, event_idx+1, sample_idx+1, ((iteration+1)<<6)+g),
//   bit 4g+j = (y_j < threshold)       src/mc/kernel/basic_event.h:112-142,
//                                      src/mc/prng/philox128.h:8-88
//   gate folds and final inversion     src/mc/kernel/gate.h:262-264, 299-362
//   ATLEAST: 8-bit count >= 8-bit k    src/mc/kernel/gate.h:506-555
//   ones += popcount(word)             src/mc/kernel/tally.h:322-360
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mcs {
namespace {

__device__ __forceinline__ void mul_wide(uint32_t a, uint32_t m, uint32_t &hi, uint32_t &lo) {
    const uint64_t p = uint64_t(a) * uint64_t(m);
    hi = uint32_t(p >> 32);
    lo = uint32_t(p);
}

// a ^ b ^ c in one VALU op (v_bitop3_b32, truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ void philox_round(uint32_t &x0, uint32_t &x1, uint32_t &x2,
                                             uint32_t &x3, uint32_t ka, uint32_t kb) {
    uint32_t ha, la, hb, lb;
    mul_wide(x0, kPhiloxM0, ha, la);
    mul_wide(x2, kPhiloxM1, hb, lb);
    x0 = xor3(hb, x1, ka);
    x1 = lb;
    x2 = xor3(ha, x3, kb);
    x3 = la;
}

// 64 Bernoulli draws for one word: Philox4x32-10 on counters
// (c0, c1, c2, c3base + g), g = 0..15, bit 4g+j = (y_j < thr).
//
// The 16 counters differ only in word 3, which enters round 1 through an
// XOR alone. So M1*c2 (round 1), M0*x0 (round 2) and M1*x2 (round 3) are the
// same for all 16 generations and are computed once per word, and the
// generation-dependent half of round 2 depends only on per-event values
// (c0, c3base, g), which are wave-uniform here and go to the scalar unit.
// That leaves 15 vector multiplies per generation plus 3 per word instead
// of 20 per generation.
__device__ __forceinline__ uint64_t bernoulli_word(uint32_t c0, uint32_t c1, uint32_t c2,
                                                   uint32_t c3base, uint32_t k0, uint32_t k1,
                                                   uint32_t thr) {
    // Round 1: (c0, c1, c2, c3) with key (k0, k1).
    uint32_t hA, lA, hB, lB;
    mul_wide(c2, kPhiloxM1, hA, lA);  // per word
    mul_wide(c0, kPhiloxM0, hB, lB);  // per event
    const uint32_t a0 = hA ^ (c1 ^ k0);
    const uint32_t a1 = lA;
    const uint32_t a2base = hB ^ k1;  // a2 = a2base ^ c3
    const uint32_t a3 = lB;
    // Round 2: the a0 half is shared by all generations.
    const uint32_t k0r1 = k0 + kPhiloxW0, k1r1 = k1 + kPhiloxW1;
    uint32_t hC, lC;
    mul_wide(a0, kPhiloxM0, hC, lC);
    const uint32_t b2 = hC ^ (a3 ^ k1r1);
    const uint32_t b3 = lC;
    // Round 3: the b2 half is shared by all generations.
    const uint32_t k0r2 = k0 + 2u * kPhiloxW0, k1r2 = k1 + 2u * kPhiloxW1;
    uint32_t hD, lD;
    mul_wide(b2, kPhiloxM1, hD, lD);

    uint32_t lo_bits = 0, hi_bits = 0;
#pragma unroll
    for (uint32_t g = 0; g < 16; ++g) {
        const uint32_t a2 = a2base ^ (c3base + g);  // per event
        uint32_t hE, lE;
        mul_wide(a2, kPhiloxM1, hE, lE);  // per event
        const uint32_t b0 = a1 ^ (hE ^ k0r1);
        const uint32_t b1 = lE;
        uint32_t hF, lF;
        mul_wide(b0, kPhiloxM0, hF, lF);
        uint32_t x0 = hD ^ (b1 ^ k0r2);
        uint32_t x1 = lD;
        uint32_t x2 = xor3(hF, b3, k1r2);
        uint32_t x3 = lF;
#pragma unroll
        for (uint32_t r = 3; r < 10; ++r)
            philox_round(x0, x1, x2, x3, k0 + r * kPhiloxW0, k1 + r * kPhiloxW1);
        const uint32_t nib = uint32_t(x0 < thr) | uint32_t(x1 < thr) << 1 |
                             uint32_t(x2 < thr) << 2 | uint32_t(x3 < thr) << 3;
        if (g < 8)
            lo_bits |= nib << (4u * g);
        else
            hi_bits |= nib << (4u * (g - 8u));
    }
    return uint64_t(hi_bits) << 32 | lo_bits;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t live_popcount(uint64_t w0, uint64_t w1, uint64_t s,
                                                  uint64_t n_bp) {
    return (s < n_bp ? __popcll(w0) : 0u) + (s + 1 < n_bp ? __popcll(w1) : 0u);
}

}  // namespace

// grid.x = n_events * n_chunks (event-major), block = 256 lanes, 2 words/lane.
__global__ __launch_bounds__(256) void sample_events(const EventDesc *__restrict__ events,
                                                     uint64_t *__restrict__ words,
                                                     unsigned long long *__restrict__ ones,
                                                     LaunchGeometry geo, uint32_t n_events,
                                                     uint32_t iteration, uint32_t k0,
                                                     uint32_t k1) {
    const uint32_t e = blockIdx.x / geo.n_chunks;
    const uint32_t chunk = blockIdx.x - e * geo.n_chunks;
    if (e >= n_events) return;
    const EventDesc d = events[e];
    const uint32_t mode = d.mode_tally & 3u;
    const bool tallied = (d.mode_tally >> 2) & 1u;
    const uint32_t c3base = (iteration + 1u) << 6;
    uint64_t *row = words + uint64_t(e) * geo.stride;
    const uint64_t base = uint64_t(chunk) * geo.chunk_words;
    uint32_t pc = 0;
    for (uint32_t off = threadIdx.x * 2u; off < geo.chunk_words; off += 512u) {
        const uint64_t s = base + off;
        uint64_t w0, w1;
        if (mode == kEventDraw) {
            w0 = bernoulli_word(d.ctr0, d.ctr1, uint32_t(s) + 1u, c3base, k0, k1, d.threshold);
            w1 = bernoulli_word(d.ctr0, d.ctr1, uint32_t(s) + 2u, c3base, k0, k1, d.threshold);
        } else {
            w0 = w1 = (mode == kEventAlways) ? ~0ull : 0ull;
        }
        *reinterpret_cast<ulonglong2 *>(row + s) = make_ulonglong2(w0, w1);
        if (tallied) pc += live_popcount(w0, w1, s, geo.n_bp);
    }
    if (tallied) {
        pc = wave_sum(pc);
        if ((threadIdx.x & 63u) == 0 && pc) atomicAdd(ones + e, (unsigned long long)pc);
    }
}

namespace {

struct Pair {
    uint64_t x, y;
};

__device__ __forceinline__ Pair load_pair(const uint64_t *__restrict__ words, uint32_t row,
                                          uint64_t stride, uint64_t s) {
    const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(words + uint64_t(row) * stride + s);
    return {v.x, v.y};
}

// Fold the inputs [positives..., negated...] with op; loads are issued four
// at a time so several rows are in flight per wave.
template <typename Op>
__device__ __forceinline__ void fold_inputs(const uint32_t *__restrict__ a, uint32_t n_in,
                                            uint32_t n_pos, const uint64_t *__restrict__ words,
                                            uint64_t stride, uint64_t s, Op &&op) {
    uint32_t i = 0;
    for (; i + 4 <= n_in; i += 4) {
        Pair v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = load_pair(words, a[i + j], stride, s);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (i + j >= n_pos) {
                v[j].x = ~v[j].x;
                v[j].y = ~v[j].y;
            }
            op(v[j]);
        }
    }
    for (; i < n_in; ++i) {
        Pair v = load_pair(words, a[i], stride, s);
        if (i >= n_pos) {
            v.x = ~v.x;
            v.y = ~v.y;
        }
        op(v);
    }
}

}  // namespace

// grid.x = n_chunks * n_gates (chunk-major: blocks running together share a
// chunk, so input rows read by several gates of the layer are reused from
// cache), block = 256 lanes, 2 words/lane.
__global__ __launch_bounds__(256) void eval_gates(const GateDesc *__restrict__ gates,
                                                  const uint32_t *__restrict__ arg_rows,
                                                  uint64_t *__restrict__ words,
                                                  unsigned long long *__restrict__ ones,
                                                  LaunchGeometry geo, uint32_t n_gates) {
    const uint32_t chunk = blockIdx.x / n_gates;
    const uint32_t gi = blockIdx.x - chunk * n_gates;
    if (chunk >= geo.n_chunks) return;
    const GateDesc d = gates[gi];
    const uint32_t n_in = d.counts & 0xFFFFu;
    const uint32_t n_pos = d.counts >> 16;
    const uint32_t type = d.type_k_tally & 0xFFu;
    const uint32_t k8 = (d.type_k_tally >> 8) & 0xFFu;
    const bool tallied = (d.type_k_tally >> 16) & 1u;
    const uint32_t *a = arg_rows + d.arg_begin;
    uint64_t *out = words + uint64_t(d.out_row) * geo.stride;
    const uint64_t base = uint64_t(chunk) * geo.chunk_words;
    uint32_t pc = 0;
    for (uint32_t off = threadIdx.x * 2u; off < geo.chunk_words; off += 512u) {
        const uint64_t s = base + off;
        Pair r;
        if (type == 2u) {  // ATLEAST: bit-sliced 8-bit counter, then count >= k8
            const uint32_t planes = n_in >= 255u ? 8u : 32u - __clz(n_in);
            Pair c[8];
#pragma unroll
            for (int b = 0; b < 8; ++b) c[b] = {0, 0};
            fold_inputs(a, n_in, n_pos, words, geo.stride, s, [&](Pair v) {
                Pair carry = v;
#pragma unroll
                for (uint32_t b = 0; b < 8; ++b) {
                    if (b < planes) {
                        const Pair t = {c[b].x & carry.x, c[b].y & carry.y};
                        c[b].x ^= carry.x;
                        c[b].y ^= carry.y;
                        carry = t;
                    }
                }
            });
            Pair gt = {0, 0}, eq = {~0ull, ~0ull};
#pragma unroll
            for (int b = 7; b >= 0; --b) {
                if ((k8 >> b) & 1u) {
                    eq.x &= c[b].x;
                    eq.y &= c[b].y;
                } else {
                    gt.x |= eq.x & c[b].x;
                    gt.y |= eq.y & c[b].y;
                    eq.x &= ~c[b].x;
                    eq.y &= ~c[b].y;
                }
            }
            r = {gt.x | eq.x, gt.y | eq.y};
        } else {
            const bool and_like = (type == 0u || type == 5u);  // AND, NAND
            r = and_like ? Pair{~0ull, ~0ull} : Pair{0, 0};
            switch (type) {
            case 0u: case 5u:  // AND, NAND
                fold_inputs(a, n_in, n_pos, words, geo.stride, s, [&](Pair v) { r.x &= v.x; r.y &= v.y; });
                break;
            case 1u: case 6u:  // OR, NOR
                fold_inputs(a, n_in, n_pos, words, geo.stride, s, [&](Pair v) { r.x |= v.x; r.y |= v.y; });
                break;
            case 3u:  // XOR
                fold_inputs(a, n_in, n_pos, words, geo.stride, s, [&](Pair v) { r.x ^= v.x; r.y ^= v.y; });
                break;
            default:  // NOT, NULL: the last input in [positives..., negated...] order
                if (n_in) {
                    r = load_pair(words, a[n_in - 1], geo.stride, s);
                    if (n_in - 1 >= n_pos) {
                        r.x = ~r.x;
                        r.y = ~r.y;
                    }
                }
                break;
            }
            if (type == 4u || type == 5u || type == 6u) {  // NOT, NAND, NOR
                r.x = ~r.x;
                r.y = ~r.y;
            }
        }
        *reinterpret_cast<ulonglong2 *>(out + s) = make_ulonglong2(r.x, r.y);
        if (tallied) pc += live_popcount(r.x, r.y, s, geo.n_bp);
    }
    if (tallied) {
        pc = wave_sum(pc);
        if ((threadIdx.x & 63u) == 0 && pc) atomicAdd(ones + d.out_row, (unsigned long long)pc);
    }
}

}  // namespace mcs

// Host-side launch wrappers (the engine's other translation units are host
// code and launch through these).
namespace mcs {

hipError_t launch_sample_events(hipStream_t stream, const EventDesc *events, uint64_t *words,
                                unsigned long long *ones, const LaunchGeometry &geo,
                                uint32_t n_events, uint32_t iteration, uint32_t k0, uint32_t k1) {
    if (n_events == 0) return hipSuccess;
    const uint64_t blocks = uint64_t(n_events) * geo.n_chunks;
    hipLaunchKernelGGL(sample_events, dim3(uint32_t(blocks)), dim3(256), 0, stream, events, words,
                       ones, geo, n_events, iteration, k0, k1);
    return hipGetLastError();
}

hipError_t launch_eval_gates(hipStream_t stream, const GateDesc *gates, const uint32_t *arg_rows,
                             uint64_t *words, unsigned long long *ones, const LaunchGeometry &geo,
                             uint32_t n_gates) {
    if (n_gates == 0) return hipSuccess;
    const uint64_t blocks = uint64_t(n_gates) * geo.n_chunks;
    hipLaunchKernelGGL(eval_gates, dim3(uint32_t(blocks)), dim3(256), 0, stream, gates, arg_rows,
                       words, ones, geo, n_gates);
    return hipGetLastError();
}

}  // namespace mcs
