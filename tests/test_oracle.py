"""Pins the CPU oracle (oracle/mc_oracle.cpp) before it is trusted as the
checker of the GPU engine.

* Philox4x32-10 (src/mc/prng/philox128.h:8-54): golden vectors produced by the
  reference's own header (tests/golden/philox_ref.json), the Random123
  known-answer vectors for arbitrary keys, and, where oracle/_ref is built,
  the reference header itself on random counters.
* Bernoulli packing and counter layout (philox128.h:59-88,
  basic_event.h:112-122): golden pack vectors; p = 0 and p = 1 special cases.
* Threshold (node.h:371-376): closed form floor(p * 2^32).
* Gates (gate.h:262-362, 506-555): per-bit Boolean definitions on random
  words, including the ATLEAST edge cases k = 0, 1, n, n + 1 and the 8-bit
  counter's wrap for n >= 256.
* Whole Monte Carlo (tally.h:322-360): exact probabilities by enumeration
  must lie in the 99.9 % interval of the estimate; the reference's core
  known answers (tests/bench_core_tests.cc) likewise.
"""
import json
import os

import numpy as np
import pytest

from oracle import binding as orc
from tests.exact import exact_probabilities
from tests.models import core_models, small_random
from mcscram_amd.pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag

HERE = os.path.dirname(os.path.abspath(__file__))
KEY = (382307844, 293830103)


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "philox_ref.json")) as f:
        return json.load(f)


def test_philox_generate_matches_reference_header_vectors(golden):
    assert tuple(golden["key"]) == KEY
    for case in golden["generate"]:
        ctr = list(case["seed"])
        ctr[3] = (ctr[3] + case["generation"]) & 0xFFFFFFFF  # generate(): x3 += generation
        assert orc.philox(ctr, KEY).tolist() == case["out"], case


# Random123 known-answer vectors for philox4x32-10 (ctr, key, output).
RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_random123_kat(ctr, key, out):
    assert tuple(orc.philox(ctr, key).tolist()) == out


def test_pack_matches_reference_header_vectors(golden):
    for c in golden["pack"]:
        w = orc.sample_word(c["pdag_idx"], c["event_idx"], c["sample_idx"], c["iteration"],
                            c["threshold"])
        assert w == c["word"], c


def test_oracle_against_reference_header_random():
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    for _ in range(300):
        pdag, ev, s, it = (int(x) for x in rng.integers(0, 2**31, 4))
        it = it % 1000 + 1
        thr = int(rng.integers(0, 2**32 + 1))
        seed = np.array([pdag + 1, ev + 1, s + 1, ((it + 1) << 6) & 0xFFFFFFFF], np.uint32)
        assert orc.sample_word(pdag, ev, s, it, thr) == ref.ref_pack_bernoulli_draws(seed.ctypes.data, thr)


def test_threshold_closed_form():
    assert orc.threshold(0.0) == 0
    assert orc.threshold(1.0) == 2**32
    assert orc.threshold(0.5) == 2**31
    assert orc.threshold(0.1) == 429496729  # 0.1 * 2^32 = 429496729.6, truncated
    rng = np.random.default_rng(3)
    for p in rng.random(200):
        assert orc.threshold(p) == int(np.floor(p * 4294967296.0))


def test_sample_word_special_probabilities():
    assert orc.sample_word(5, 0, 0, 1, 0) == 0
    assert orc.sample_word(5, 0, 0, 1, 2**32) == (1 << 64) - 1


def test_sample_word_bit_layout():
    # bit 4g + j of a word is draw j of generation g: counter word 3 =
    # ((iteration + 1) << 6) + g (philox128.h:64-86, basic_event.h:112-122).
    pdag, ev, s, it = 17, 3, 42, 2
    thr = orc.threshold(0.37)
    w = orc.sample_word(pdag, ev, s, it, thr)
    for g in range(16):
        y = orc.philox([pdag + 1, ev + 1, s + 1, ((it + 1) << 6) + g], KEY)
        for j in range(4):
            assert ((w >> (4 * g + j)) & 1) == int(int(y[j]) < thr)


def _bits(words, n_bits=64):
    return [[(int(w) >> b) & 1 for w in words] for b in range(n_bits)]


@pytest.mark.parametrize("conn", [AND, OR, XOR, NAND, NOR])
def test_gate_words_truth_table(conn):
    rng = np.random.default_rng(conn + 11)
    for n in (1, 2, 3, 5, 8):
        for n_neg in range(0, n + 1):
            vals = rng.integers(0, 2**64, size=n, dtype=np.uint64)
            got = orc.gate_word(conn, 0, vals, n_neg)
            for b, col in enumerate(_bits(vals)):
                lits = [v if i < n - n_neg else 1 - v for i, v in enumerate(col)]
                ref = {AND: all(lits), OR: any(lits), XOR: sum(lits) % 2 == 1,
                       NAND: not all(lits), NOR: not any(lits)}[conn]
                assert ((got >> b) & 1) == int(ref)


@pytest.mark.parametrize("conn", [NOT, NULL])
def test_gate_words_single_input(conn):
    rng = np.random.default_rng(5)
    v = rng.integers(0, 2**64, size=1, dtype=np.uint64)
    for n_neg in (0, 1):
        lit = int(v[0]) if n_neg == 0 else (~int(v[0])) & (2**64 - 1)
        want = lit if conn == NULL else (~lit) & (2**64 - 1)
        assert orc.gate_word(conn, 0, v, n_neg) == want


def test_atleast_truth_table_and_edges():
    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 6, 9, 16):
        for k in sorted({0, 1, max(n // 2, 1), n, n + 1}):
            for n_neg in (0, n // 3):
                vals = rng.integers(0, 2**64, size=n, dtype=np.uint64)
                got = orc.gate_word(ATLEAST, k, vals, n_neg)
                for b, col in enumerate(_bits(vals)):
                    lits = [v if i < n - n_neg else 1 - v for i, v in enumerate(col)]
                    assert ((got >> b) & 1) == int(sum(lits) >= k)
    # k = 0 is always true; k > n never.
    v = rng.integers(0, 2**64, size=4, dtype=np.uint64)
    assert orc.gate_word(ATLEAST, 0, v, 0) == 2**64 - 1
    assert orc.gate_word(ATLEAST, 5, v, 0) == 0


def test_atleast_eight_bit_counter_wraps():
    # The reference counts true inputs in a uint8 (gate.h:526-541): with 257
    # all-true inputs the count wraps to 1, so k = 2 fails while k = 1 holds.
    ones = np.full(257, np.uint64(2**64 - 1), dtype=np.uint64)
    assert orc.gate_word(ATLEAST, 2, ones, 0) == 0
    assert orc.gate_word(ATLEAST, 1, ones, 0) == 2**64 - 1


def _ci_contains(ones, total, exact, z=3.2905):  # 99.9 % two-sided
    mean = ones / total
    se = max(np.sqrt(mean * (1 - mean) / total), 1.0 / total)
    return abs(mean - exact) <= z * se


@pytest.mark.parametrize("name", sorted(core_models()))
def test_core_known_answers(name):
    pd, exact = core_models()[name]
    assert abs(exact_probabilities(pd)[pd.root] - exact) < 1e-12
    n_bp = 1 << 14
    ones, _ = orc.run(pd, n_bp)
    assert _ci_contains(int(ones[0]), n_bp * 64, exact), (name, ones[0] / (n_bp * 64), exact)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_small_trees_against_enumeration(seed):
    pd = small_random(seed)
    exact = exact_probabilities(pd)
    tally = [int(g) for g in pd.gate_index] + [int(v) for v in pd.var_index]
    n_bp = 1 << 13
    ones, _ = orc.run(pd, n_bp, tally=tally)
    for idx, o in zip(tally, ones):
        assert _ci_contains(int(o), n_bp * 64, exact[idx]), (seed, idx)


def test_run_words_consistent_with_gate_words():
    # Every gate word of a pass equals the gate function of its input words.
    pd = small_random(11)
    ones, words = orc.run(pd, 64, tally=[pd.root], want_words=True)
    row = {int(v): s for s, v in enumerate(pd.var_index)}
    row.update({int(g): pd.n_vars + s for s, g in enumerate(pd.gate_index)})
    for s, g in enumerate(pd.gate_index):
        args = pd.gate_args(s)
        vars_first = [a for a in args if abs(a) in set(pd.var_index.tolist())] + \
                     [a for a in args if abs(a) not in set(pd.var_index.tolist())]
        pos = [a for a in vars_first if a > 0]
        neg = [a for a in vars_first if a < 0]
        for w in range(64):
            vals = np.array([words[row[abs(a)], w] for a in pos + neg], dtype=np.uint64)
            assert orc.gate_word(int(pd.gate_type[s]), int(pd.gate_k[s]), vals, len(neg)) == \
                words[row[int(g)], w]
    assert int(ones[0]) == sum(bin(int(x)).count("1") for x in words[row[pd.root]])


def test_iterations_are_independent_streams():
    # Pass i uses counter word 3 = ((i + 1) << 6) + g: passes 1..3 in one call
    # equal the sum of three single-pass calls at iterations 1, 2, 3.
    pd = small_random(5)
    tally = [int(g) for g in pd.gate_index]
    both, _ = orc.run(pd, 256, first_iter=1, n_iters=3, tally=tally)
    parts = sum(orc.run(pd, 256, first_iter=i, n_iters=1, tally=tally)[0] for i in (1, 2, 3))
    assert np.array_equal(both, parts)


def test_point_estimates_closed_form():
    out = orc.point_estimates(250, 1000)
    se = np.sqrt(0.25 * 0.75 / 1000)
    np.testing.assert_allclose(out, [0.25, se, 0.25 - 1.959963984540054 * se,
                                     0.25 + 1.959963984540054 * se,
                                     0.25 - 2.5758293035489004 * se,
                                     0.25 + 2.5758293035489004 * se], rtol=0, atol=1e-15)
    assert orc.point_estimates(0, 64)[4] == 0.0  # clamped at 0
    assert orc.point_estimates(64, 64)[5] == 1.0  # clamped at 1


def test_avx512_oracle_build_equals_the_generic_one():
    # liboracle_v4.so draws the 16 generations of a word as vector lanes;
    # its words must equal the generic build's bit for bit, thresholds 0
    # and 2^32 included.
    import ctypes as C
    import os
    if not (orc.host_has_avx512() and os.path.exists(os.path.join(orc._HERE, "liboracle_v4.so"))):
        pytest.skip("no AVX-512 host or no liboracle_v4.so")
    libs = [C.CDLL(os.path.join(orc._HERE, n)) for n in ("liboracle.so", "liboracle_v4.so")]
    for L in libs:
        L.orc_sample_word.argtypes = [C.c_uint32] * 4 + [C.c_uint64]
        L.orc_sample_word.restype = C.c_uint64
    rng = np.random.default_rng(17)
    thresholds = [0, 1, 1 << 20, (1 << 31) + 5, (1 << 32) - 1, 1 << 32]
    for _ in range(3000):
        a = [int(v) for v in rng.integers(0, 1 << 32, size=4, dtype=np.uint64)]
        thr = thresholds[int(rng.integers(0, len(thresholds)))] if rng.random() < 0.3 \
            else int(rng.integers(0, (1 << 32) + 1))
        assert libs[0].orc_sample_word(*a, thr) == libs[1].orc_sample_word(*a, thr), (a, thr)
