"""Lognormal parameter samples drawn on the GPU (mcs_lognormal_samples).

CPU: the numpy oracle's Philox4x32-10 against the Random123 known answers,
the reference header's own vectors and the C oracle on random counters and
keys; its Box-Muller transform at uniforms where the closed form is known
(u1 = 1/2, u2 in {0, 1/4, 1/2, 3/4}), which also fixes which output words
feed which uniform; the normal distribution of 200,000 draws.
GPU: the kernel against the oracle (z within one FP32 rounding, p within
1e-12 relative: both sides evaluate log, sqrt, sin, cos and exp in FP64 with
their own libraries), rows independent of how calls split them, row ids
past 2^32, odd variable counts; invalid input is refused before any device
call (a CPU test).
"""
import json
import os

import numpy as np
import pytest
from scipy import stats

from mcscram_amd import ua
from oracle import binding as orc
from oracle import lognormal_ref as ref
from tests.test_oracle import KEY, RANDOM123_KAT

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_numpy_philox_random123_kat(ctr, key, out):
    y = ref.philox4x32_10([np.uint32(c) for c in ctr], key)
    assert tuple(int(v) for v in y) == out


def test_numpy_philox_reference_header_vectors():
    with open(os.path.join(HERE, "golden", "philox_ref.json")) as f:
        golden = json.load(f)
    assert tuple(golden["key"]) == KEY
    for case in golden["generate"]:
        ctr = list(case["seed"])
        ctr[3] = (ctr[3] + case["generation"]) & 0xFFFFFFFF
        y = ref.philox4x32_10([np.uint32(c) for c in ctr], KEY)
        assert [int(v) for v in y] == case["out"], case


def test_numpy_philox_matches_c_oracle_on_random_counters_and_keys():
    rng = np.random.default_rng(5)
    ctr = rng.integers(0, 1 << 32, size=(4, 300), dtype=np.uint64)
    keys = rng.integers(0, 1 << 32, size=(300, 2), dtype=np.uint64)
    for i in range(300):
        key = (int(keys[i, 0]), int(keys[i, 1]))
        y = ref.philox4x32_10([ctr[w, i] for w in range(4)], key)
        want = orc.philox([int(ctr[w, i]) for w in range(4)], key).tolist()
        assert [int(v) for v in y] == want


def test_box_muller_closed_form_points():
    # (y0, y1) = (2^31, 0): b1 = 2^52, u1 = (2^52 + 1/2) / 2^53 (1/2 to 1e-16);
    # (y2, y3) = (q 2^30, 0): u2 = q / 4.
    y0 = np.array([1 << 31] * 4, dtype=np.uint32)
    y1 = np.zeros(4, dtype=np.uint32)
    y2 = np.array([0, 1 << 30, 2 << 30, 3 << 30], dtype=np.uint32)
    y3 = np.zeros(4, dtype=np.uint32)
    ze, zo = ref.box_muller(y0, y1, y2, y3)
    r = np.sqrt(2.0 * np.log(2.0))
    np.testing.assert_allclose(ze, [r, 0.0, -r, 0.0], atol=1e-12)
    np.testing.assert_allclose(zo, [0.0, r, 0.0, -r], atol=1e-12)
    # The low word of a pair carries the low bits: (0, 2^31) moves u1 by 2^-22
    # of its range only, (2^31, 0) to 1/2.
    ze2, _ = ref.box_muller(np.zeros(1, np.uint32), np.array([1 << 31], np.uint32),
                            np.zeros(1, np.uint32), np.zeros(1, np.uint32))
    u1 = ((1 << 31) >> 11) / 2.0 ** 53 + 0.5 / 2.0 ** 53
    np.testing.assert_allclose(ze2, [np.sqrt(-2.0 * np.log(u1))], rtol=1e-12)


def test_oracle_draws_are_standard_normal():
    z = ref.normals(20, seed=77, first_row=0, n_rows=10_000)  # 200,000 draws
    flat = z.ravel()
    assert abs(flat.mean()) < 5.0 / np.sqrt(flat.size)
    assert abs(flat.var() - 1.0) < 5.0 * np.sqrt(2.0 / flat.size)
    assert stats.kstest(flat, "norm").pvalue > 1e-4
    # the two members of a Box-Muller pair, and neighbouring rows, uncorrelated
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.05
    assert abs(np.corrcoef(z[:-1, 3], z[1:, 3])[0, 1]) < 0.05
    # a different seed is a different stream
    assert not np.allclose(ref.normals(20, seed=78, first_row=0, n_rows=4), z[:4])


def test_oracle_p_is_the_clamped_lognormal():
    mu = np.array([np.log(1e-3), 0.0, np.log(0.5)])
    sigma = np.array([1.0, 1.0, 0.0])
    z, p = ref.lognormal_samples(mu, sigma, seed=3, first_row=0, n_rows=500)
    np.testing.assert_allclose(p[:, 0], 1e-3 * np.exp(z[:, 0]), rtol=1e-13)
    assert np.all(p[:, 1] <= 1.0) and np.any(p[:, 1] == 1.0)
    np.testing.assert_array_equal(p[:, 1][z[:, 1] >= 0], 1.0)
    np.testing.assert_allclose(p[:, 2], 0.5, rtol=1e-15)


def _model(n_vars, seed=2):
    rng = np.random.default_rng(seed)
    return ua.LognormalModel(10.0 ** rng.uniform(-5, -1, n_vars), rng.uniform(3, 10, n_vars))


def _check_against_oracle(model, seed, first_row, n_rows):
    p, z = model.draw(n_rows, seed, first_row)
    zr, pr = ref.lognormal_samples(model.mu, model.sigma, seed, first_row, n_rows)
    assert z.dtype == np.float32 and p.dtype == np.float64
    np.testing.assert_allclose(z, zr.astype(np.float32), rtol=0, atol=1e-6 * max(1.0, np.abs(zr).max()))
    np.testing.assert_allclose(p, pr, rtol=1e-12, atol=0)
    return p, z


@pytest.mark.gpu
@pytest.mark.parametrize("n_vars", [1, 7, 2000])
def test_gpu_draws_match_the_oracle(gpu, n_vars):
    _check_against_oracle(_model(n_vars), seed=1234, first_row=0, n_rows=257)


@pytest.mark.gpu
def test_gpu_draws_past_row_2_to_the_32(gpu):
    _check_against_oracle(_model(33), seed=(7 << 32) | 99, first_row=(1 << 32) - 5, n_rows=11)


@pytest.mark.gpu
def test_gpu_rows_do_not_depend_on_the_split(gpu):
    m = _model(301)
    p, z = m.draw(1000, 5)
    parts = list(m.draw_chunks(1000, 5, 333))
    assert [a for a, _, _ in parts] == [0, 333, 666, 999]
    np.testing.assert_array_equal(np.concatenate([q for _, q, _ in parts]), p)
    np.testing.assert_array_equal(np.concatenate([y for _, _, y in parts]), z)


@pytest.mark.gpu
def test_gpu_draws_at_the_uncertainty_workload_size(gpu):
    _, model = ua.uncertainty_tree()
    p, z = model.draw(16384, 1234)
    flat = z.astype(np.float64).ravel()
    assert abs(flat.mean()) < 5.0 / np.sqrt(flat.size)
    assert abs(flat.var() - 1.0) < 5.0 * np.sqrt(2.0 / flat.size)
    np.testing.assert_allclose(np.median(np.log(p), axis=0), model.mu, atol=0.1 * model.sigma.max())
    # spot rows against the oracle
    for r in (0, 8191, 16383):
        zr, pr = ref.lognormal_samples(model.mu, model.sigma, 1234, r, 1)
        np.testing.assert_allclose(p[r], pr[0], rtol=1e-12)


def test_draws_reject_invalid_input_before_touching_a_device():
    from mcscram_amd.engine import EngineError
    bad = ua.LognormalModel(np.array([1e-3, 1e-2]), np.array([3.0, 0.5]))  # EF < 1: sigma < 0
    with pytest.raises(EngineError, match="sigma"):
        bad.draw(4, 1)
    inf = ua.LognormalModel(np.array([1e-3, 0.0]), np.array([3.0, 3.0]))  # ln 0: mu = -inf
    with np.errstate(divide="ignore"), pytest.raises(EngineError, match="finite mu"):
        inf.draw(4, 1)
    p, z = _model(5).draw(0, 1)  # no rows: nothing to draw
    assert p.shape == (0, 5) and z.shape == (0, 5)
