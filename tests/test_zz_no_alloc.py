"""The allocation-free entry point, mcs_probability_analysis_noalloc, against
mcs_probability_analysis through the C ABI: identical tallies, pass counts
and root estimates on two calls in a row (the second on a smaller tree and
its own stream), a clear error for an engine larger than the fixed scratch,
and no change in the device's free memory across calls.
"""
import ctypes as C

import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd.engine import EngineError
from tests.models import core_models

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _free_bytes(hip):
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _same(a, b):
    pa, ta, na = a
    pb, tb, nb = b
    assert na == nb
    assert pa == pb
    assert [(t.index, t.num_one_bits, t.total_bits) for t in ta] == \
        [(t.index, t.num_one_bits, t.total_bits) for t in tb]
    assert [t.ci for t in ta] == [t.ci for t in tb]


def test_noalloc_matches_the_allocating_entry_point_twice(gpu):
    hip = _hip()
    big = mc.small_tree(1)  # 200 events, 120 gates
    gates = [int(g) for g in big.gate_index]
    fixed = mc.Settings(num_trials=1 << 21, early_stop=False)
    # First call: the small workload's shape, 16,384 words per pass, two
    # passes, every gate observed (about 40 MiB of words: one event buffer).
    want1 = mc.probability_analysis(big, fixed, (1, 16384), gates)
    got1 = mc.probability_analysis(big, fixed, (1, 16384), gates, no_alloc=True)
    _same(got1, want1)
    assert got1[2] == 2 and got1[1][0].total_bits == 1 << 21
    # Second call, smaller: a three-event model with early stop, on a stream
    # of the caller's. The scratch keeps the first call's contents; the
    # counters must start again from zero.
    pd, exact = core_models()["abc"]
    es = mc.Settings(num_trials=0, burn_in_trials=1 << 18, ci_rel_margin=0.01)
    want2 = mc.probability_analysis(pd, es, 1 << 12)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        got2 = mc.probability_analysis(pd, es, 1 << 12, no_alloc=True, stream=stream.value)
    finally:
        hip.hipStreamDestroy(stream)
    _same(got2, want2)
    assert got2[2] == 2 and abs(got2[0] - exact) <= 3.2905 * got2[1][0].std_err
    # Repeated calls change nothing on the device: no memory is taken.
    before = _free_bytes(hip)
    for _ in range(3):
        again = mc.probability_analysis(big, fixed, (1, 16384), gates, no_alloc=True)
        _same(again, want1)
    assert _free_bytes(hip) == before


def test_noalloc_rejects_an_engine_beyond_the_scratch(gpu):
    big = mc.small_tree(1)
    # 2^18 words per pass: 320 rows of 2 MiB, far beyond the 64 MiB scratch.
    with pytest.raises(EngineError) as e:
        mc.probability_analysis(big, mc.Settings(num_trials=1 << 24, early_stop=False),
                                1 << 18, no_alloc=True)
    assert e.value.status == 4
    assert "no-allocation scratch" in str(e.value) and str(64 << 20) in str(e.value)
    # The scratch is still usable afterwards.
    pd, _ = core_models()["atleast"]
    s = mc.Settings(num_trials=1 << 20, early_stop=False)
    _same(mc.probability_analysis(pd, s, 1 << 12, no_alloc=True),
          mc.probability_analysis(pd, s, 1 << 12))
    assert mc.NOALLOC_SCRATCH_BYTES == 64 << 20
