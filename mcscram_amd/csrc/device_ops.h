// Device functions shared by the engine's kernels: the Philox4x32-10
// Bernoulli word, wave reductions and the ATLEAST plane comparison.
//
// Reference semantics restated here:
//   Philox4x32-10 round and key schedule   src/mc/prng/philox128.h:8-54,
//                                          src/mc/prng/state128.h:21-26
//   counter (pdag_idx+1, event_idx+1, sample_idx+1, ((iteration+1)<<6)+g),
//   bit 4g+j = (y_j < threshold)           src/mc/kernel/basic_event.h:112-142,
//                                          src/mc/prng/philox128.h:59-88
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace mcs {

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

// acc * 2 + (x < thr): the subtraction's borrow is the comparison, and an
// add with carry shifts it in, two VALU instructions per draw. (Written as
// C++, with or without the carry builtins, the compiler emits a compare, a
// select and a shift-or instead.) thr is wave-uniform.
__device__ __forceinline__ uint32_t shift_in_below(uint32_t acc, uint32_t x, uint32_t thr) {
    uint32_t diff;
    asm("v_sub_co_u32 %1, vcc, %2, %3\n\tv_addc_co_u32 %0, vcc, %0, %0, vcc"
        : "+v"(acc), "=&v"(diff)
        : "v"(x), "s"(thr)
        : "vcc");
    return acc;
}

// 64 Bernoulli draws for one word: Philox4x32-10 on counters
// (c0, c1, c2, c3base + g), g = 0..15, bit 4g+j = (y_j < thr).
//
// The 16 counters differ only in word 3, which enters round 1 through an
// XOR alone. So M1*c2 (round 1), M0*x0 (round 2) and M1*x2 (round 3) are the
// same for all 16 generations and are computed once per word, and the
// generation-dependent half of round 2 depends only on per-event values
// (c0, c3base, g), which go to the scalar unit when they are wave-uniform.
//
// Packing, two ways:
//  * kDense = false: the four draws of a generation hit only if their
//    minimum is below the threshold; when no lane of the wave hits (most
//    generations at the small probabilities of fault trees) the compares
//    and shifts are skipped by one wave-uniform branch.
//  * kDense = true: no branch; generations run from 15 down to 0 and each
//    draw's comparison is shifted into the word (subtract-with-borrow +
//    add-with-carry), for waves where hits are frequent and the branch
//    would always be taken.
template <bool kDense>
__device__ __forceinline__ uint64_t bernoulli_word_t(uint32_t c0, uint32_t c1, uint32_t c2,
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

    // Outputs (x0..x3) of generation g.
    auto draw = [&](uint32_t g, uint32_t &x0, uint32_t &x1, uint32_t &x2, uint32_t &x3) {
        const uint32_t a2 = a2base ^ (c3base + g);  // per event
        uint32_t hE, lE;
        mul_wide(a2, kPhiloxM1, hE, lE);  // per event
        const uint32_t b0 = a1 ^ (hE ^ k0r1);
        const uint32_t b1 = lE;
        uint32_t hF, lF;
        mul_wide(b0, kPhiloxM0, hF, lF);
        x0 = hD ^ (b1 ^ k0r2);
        x1 = lD;
        x2 = xor3(hF, b3, k1r2);
        x3 = lF;
#pragma unroll
        for (uint32_t r = 3; r < 10; ++r)
            philox_round(x0, x1, x2, x3, k0 + r * kPhiloxW0, k1 + r * kPhiloxW1);
    };

    uint32_t lo_bits = 0, hi_bits = 0;
    if (kDense) {
#pragma unroll
        for (int g = 15; g >= 0; --g) {
            uint32_t x0, x1, x2, x3;
            draw(uint32_t(g), x0, x1, x2, x3);
            uint32_t &acc = g >= 8 ? hi_bits : lo_bits;
            acc = shift_in_below(acc, x3, thr);
            acc = shift_in_below(acc, x2, thr);
            acc = shift_in_below(acc, x1, thr);
            acc = shift_in_below(acc, x0, thr);
        }
    } else {
#pragma unroll
        for (uint32_t g = 0; g < 16; ++g) {
            uint32_t x0, x1, x2, x3;
            draw(g, x0, x1, x2, x3);
            const uint32_t m = min(min(x0, x1), min(x2, x3));
            if (__builtin_amdgcn_ballot_w64(m < thr) != 0) {
                const uint32_t nib = uint32_t(x0 < thr) | uint32_t(x1 < thr) << 1 |
                                     uint32_t(x2 < thr) << 2 | uint32_t(x3 < thr) << 3;
                if (g < 8)
                    lo_bits |= nib << (4u * g);
                else
                    hi_bits |= nib << (4u * (g - 8u));
            }
        }
    }
    return uint64_t(hi_bits) << 32 | lo_bits;
}

// VALU instructions per Philox block of a sampler-only launch (rocprofv3
// SQ_INSTS_VALU, 2,048 events at one p, stores and summaries included): the
// branch-free packing 41.3 at any p; the branching one 36.4 at p = 1e-5,
// 38.8 at 1e-3, 40.7 at 2e-3, 42.2 at 3e-3, 44.9 at 6e-3 and 47.2 at 0.5
// (the chance that some lane of the wave hits in a generation grows as
// 1 - (1 - p)^256). They break even near p = 2.4e-3: thresholds at or above
// 0.625 x 2^24 take the branch-free packing.
constexpr uint32_t kDenseThreshold = (1u << 23) + (1u << 21);

// The Bernoulli word with the packing chosen per wave.
__device__ __forceinline__ uint64_t bernoulli_word(uint32_t c0, uint32_t c1, uint32_t c2,
                                                   uint32_t c3base, uint32_t k0, uint32_t k1,
                                                   uint32_t thr) {
    if (__builtin_amdgcn_ballot_w64(thr >= kDenseThreshold) != 0)
        return bernoulli_word_t<true>(c0, c1, c2, c3base, k0, k1, thr);
    return bernoulli_word_t<false>(c0, c1, c2, c3base, k0, k1, thr);
}

// Sum of v over a fully active wave64, returned as a wave-uniform value.
// DPP adds inside each 16-lane row (the inputs shifted right by 1, 2 and 3
// lanes, then the partial sums by 4 lanes into lanes 4-15 and by 8 lanes
// into lanes 8-15) leave each row's total in its lane 15; row broadcasts of
// lane 15 into rows 1 and 3, then of lane 31 into rows 2 and 3, leave the
// wave's total in lane 63. Seven VALU adds and no LDS round trips.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    uint32_t s = v;
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, true));  // row_shr:1
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, true));  // row_shr:2
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x113, 0xf, 0xf, true));  // row_shr:3
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(s), 0x114, 0xf, 0xe, true));  // row_shr:4
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(s), 0x118, 0xf, 0xc, true));  // row_shr:8
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(s), 0x142, 0xa, 0xf, true));  // row_bcast:15
    s += uint32_t(__builtin_amdgcn_update_dpp(0, int(s), 0x143, 0xc, 0xf, true));  // row_bcast:31
    return uint32_t(__builtin_amdgcn_readlane(int(s), 63));
}

// ATLEAST: a bit-sliced counter (carry-save planes, plane b = bit b of each
// trial's count of true inputs) folded over the inputs, then the per-bit
// comparison count >= k8 on the planes.
template <typename W, int P>
__device__ __forceinline__ W atleast_compare(const W (&c)[P], uint32_t k8) {
    W gt = W(0), eq = ~W(0);
#pragma unroll
    for (int b = P - 1; b >= 0; --b) {
        if ((k8 >> b) & 1u) {
            eq &= c[b];
        } else {
            gt |= eq & c[b];
            eq &= ~c[b];
        }
    }
    return gt | eq;
}

}  // namespace mcs
