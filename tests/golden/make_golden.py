"""Writes tests/golden/philox_ref.json from the reference's own Philox code.

oracle/Makefile builds the reference's src/mc/prng/philox128.h behind a C
shim into oracle/_ref/libref_philox.so (that happens only where the
reference tree exists). This script calls it on a fixed set of counters and
thresholds and stores inputs and outputs as a fixture, so the tests can pin
the oracle and the GPU sampler to the reference's outputs anywhere.

Run: make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from oracle import binding  # noqa: E402


def main():
    ref = binding.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_philox.so not built (make -C oracle with the reference tree)")
    rng = np.random.default_rng(12345)
    gen_cases, pack_cases = [], []
    # generate(): counters with edge values and random ones, all 16 generations.
    seeds = [[0, 0, 0, 0], [1, 1, 1, 128], [0xFFFFFFFF] * 4, [3, 7, 1, 0xFFFFFFC0]]
    seeds += rng.integers(0, 2**32, size=(12, 4), dtype=np.uint64).tolist()
    for s in seeds:
        s = [int(v) for v in s]
        for g in (0, 1, 7, 15, 63):
            out = np.zeros(4, np.uint32)
            ref.ref_philox_generate(np.array(s, np.uint32).ctypes.data, g, out.ctypes.data)
            gen_cases.append({"seed": s, "generation": g, "out": [int(v) for v in out]})
    # pack_bernoulli_draws(): the sampler's counters (pdag+1, event+1, sample+1,
    # (iteration+1)<<6) and thresholds including 0, 1, 2^31, 2^32 - 1, 2^32.
    thresholds = [0, 1, 429496, 42949672, 2**31, 2**32 - 1, 2**32]
    for pdag in (2, 3, 201, 25001):
        for ev in (0, 5, 199):
            for sidx in (0, 1, 16383, 262143):
                for it in (1, 2, 255):
                    seed = [pdag + 1, ev + 1, sidx + 1, ((it + 1) << 6) & 0xFFFFFFFF]
                    thr = thresholds[(pdag + ev + sidx + it) % len(thresholds)]
                    word = ref.ref_pack_bernoulli_draws(np.array(seed, np.uint32).ctypes.data, thr)
                    pack_cases.append({"pdag_idx": pdag, "event_idx": ev, "sample_idx": sidx,
                                       "iteration": it, "threshold": thr, "word": int(word)})
    with open(os.path.join(HERE, "philox_ref.json"), "w") as f:
        json.dump({"source": "mcSCRAM src/mc/prng/philox128.h built by oracle/Makefile",
                   "key": [382307844, 293830103], "generate": gen_cases, "pack": pack_cases},
                  f, indent=0)
    print(len(gen_cases), "generate cases,", len(pack_cases), "pack cases")


if __name__ == "__main__":
    main()
