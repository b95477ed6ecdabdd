// Test-only shim that exposes the reference's own Philox header through a C
// ABI so the oracle's restatement can be checked against it. It contains no
// reference code: the reference header is included from its location in the
// reference tree at build time (see oracle/Makefile) and the result lands in
// oracle/_ref/, which is never committed.
#include <cstdint>
#include "mc/prng/philox128.h"

namespace ph = scram::mc::prng;

extern "C" {

// philox::generate with the reference's fixed key: counter = seed with
// x[3] += generation (src/mc/prng/philox128.h:37-54).
void ref_philox_generate(const uint32_t seed[4], uint32_t generation, uint32_t out[4]) {
    ph::state128 s{{seed[0], seed[1], seed[2], seed[3]}};
    ph::state128 r{};
    ph::philox::generate(&s, &r, static_cast<uint8_t>(generation));
    for (int i = 0; i < 4; ++i) out[i] = r.x[i];
}

// philox::pack_bernoulli_draws<uint64_t> (src/mc/prng/philox128.h:76-88).
uint64_t ref_pack_bernoulli_draws(const uint32_t seed[4], uint64_t threshold) {
    ph::state128 s{{seed[0], seed[1], seed[2], seed[3]}};
    return ph::philox::pack_bernoulli_draws<uint64_t>(s, threshold);
}

}
