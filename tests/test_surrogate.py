"""Surrogate network: the CPU oracle pinned by finite differences, then the
GPU matrix-core products and gradients against the oracle, and training
against plain Monte Carlo targets.
"""
import numpy as np
import pytest

from oracle import surrogate_ref as ref


def _params(n_in, h, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    n = n_in * h + 2 * h * h + 4 * h + 1
    p = rng.standard_normal(n) * scale / np.sqrt(h)
    return p


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 0.0], dtype=np.float32)
    r = ref.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # halfway: to even (mantissa 0)
    assert r[2] == np.float32(1.0 + 2**-6)  # halfway: to even (up)
    assert r[3] == np.float32(1.0 + 2**-7)  # above halfway: up
    assert r[4] == -2.5 and r[5] == 0.0


def test_oracle_gradient_matches_finite_differences():
    # float64, no rounding: central differences of the forward loss.
    n_in, h = 5, 4
    rng = np.random.default_rng(3)
    x = rng.standard_normal((7, n_in))
    y = rng.standard_normal(7)
    p = _params(n_in, h, 1, scale=2.0)
    loss, g = ref.loss_and_grad(p, x, y, n_in, h, bf16_mode=False)
    eps = 1e-6
    for i in range(p.shape[0]):
        q = p.copy()
        q[i] += eps
        lp, _ = ref.loss_and_grad(q, x, y, n_in, h, bf16_mode=False)
        q[i] -= 2 * eps
        lm, _ = ref.loss_and_grad(q, x, y, n_in, h, bf16_mode=False)
        assert abs((lp - lm) / (2 * eps) - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_oracle_prediction_consistent_with_loss():
    n_in, h = 6, 8
    rng = np.random.default_rng(4)
    x = rng.standard_normal((9, n_in))
    y = rng.standard_normal(9)
    p = _params(n_in, h, 2)
    yhat = ref.predict(p, x, n_in, h, bf16_mode=False)
    loss, _ = ref.loss_and_grad(p, x, y, n_in, h, bf16_mode=False)
    assert abs(loss - np.mean((yhat - y) ** 2)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("shape,ta,tb", [((64, 64, 32), False, False), ((128, 192, 96), False, False),
                                          ((64, 128, 64), True, False), ((128, 64, 64), False, True),
                                          ((192, 64, 256), True, True),
                                          # split-K: few output tiles, long K (even and
                                          # uneven splits, the at-scale first layer)
                                          ((64, 512, 4096), False, False),
                                          ((128, 64, 1184), False, True),
                                          ((256, 512, 25024), False, False)])
def test_mfma_gemm_against_numpy(gpu, shape, ta, tb):
    from mcscram_amd.surrogate import gemm_test
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    b = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    c = gemm_test(a, b, ta, tb)
    A = ref.bf16(a).astype(np.float64)
    B = ref.bf16(b).astype(np.float64)
    want = (A.T if ta else A) @ (B.T if tb else B)
    # FP32 accumulation of exact BF16 products: error ~ K * 2^-24 * |terms|.
    tol = 4 * K * 2.0**-24 * (np.abs(A.T if ta else A) @ np.abs(B.T if tb else B)) + 1e-6
    assert np.all(np.abs(c - want) <= tol), np.max(np.abs(c - want) - tol)


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(gpu):
    from mcscram_amd.surrogate import Mlp
    n_in, h = 100, 64
    net = Mlp(n_in, h, seed=3)
    p = net.params()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((128, n_in)).astype(np.float32)
    y = rng.standard_normal(128).astype(np.float32)
    loss, g = net.loss_grad(x, y)
    rloss, rg = ref.loss_and_grad(p, x, y, n_in, h, bf16_mode=True)
    assert abs(loss - rloss) <= 1e-4 * max(1.0, rloss)
    # Accumulation order differs, so BF16 roundings of backpropagated values
    # can differ by one unit now and then: compare norm-wise, per tensor.
    sizes = [n_in * h, h, h * h, h, h * h, h, h, 1]
    o = 0
    for name, n in zip(("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"), sizes):
        d = np.linalg.norm(g[o:o + n] - rg[o:o + n])
        assert d <= 2e-2 * np.linalg.norm(rg[o:o + n]) + 1e-7, (name, d, np.linalg.norm(rg[o:o + n]))
        o += n


@pytest.mark.gpu
def test_set_params_roundtrip_and_prediction(gpu):
    from mcscram_amd.surrogate import Mlp
    n_in, h = 37, 64
    net = Mlp(n_in, h, seed=9)
    p = _params(n_in, h, 6).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    x = np.random.default_rng(1).standard_normal((70, n_in)).astype(np.float32)
    got = net.predict(x)
    want = ref.predict(p, x, n_in, h, bf16_mode=True)
    assert np.allclose(got, want, rtol=1e-2, atol=1e-3)


@pytest.mark.gpu
def test_training_fits_a_smooth_function(gpu):
    from mcscram_amd.surrogate import Mlp
    n_in, h = 64, 128
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4096, n_in)).astype(np.float32)
    w = rng.standard_normal(n_in) / np.sqrt(n_in)
    y = (np.tanh(x @ w) + 0.3 * (x[:, 0] * x[:, 1])).astype(np.float32)
    net = Mlp(n_in, h, seed=4)
    losses = net.fit(x, y, epochs=20, batch=256, lr=2e-3)
    assert losses[-1] < 0.2 * losses[0], losses
    xt = rng.standard_normal((512, n_in)).astype(np.float32)
    yt = np.tanh(xt @ w) + 0.3 * (xt[:, 0] * xt[:, 1])
    mse = float(np.mean((net.predict(xt) - yt) ** 2))
    assert mse < 0.5 * float(np.var(yt)), mse  # held-out R^2 above 0.5


def test_lognormal_input_map_is_probability_over_q99():
    # min(cap, exp(sigma z - Z99 sigma)) must equal min(1, exp(mu + sigma z))
    # / exp(mu + Z99 sigma), clamped events included.
    from mcscram_amd import screening as scr
    from mcscram_amd import ua
    m = ua.LognormalModel(np.array([1e-5, 1e-3, 0.3, 0.9]), np.array([3.0, 10.0, 10.0, 5.0]))
    p, z = m.sample(5000, seed=4)
    assert (p == 1.0).any()
    scale, shift, cap = scr.lognormal_input_map(m)
    u = np.minimum(cap[None, :], np.exp(scale[None, :] * z + shift[None, :]))
    q99 = np.exp(m.mu + scr.Z99 * m.sigma)
    assert np.allclose(u, p / q99[None, :], rtol=1e-12)
    # About 1 % of the unclamped draws exceed their 99th percentile.
    assert abs((u[:, :2] > 1.0).mean() - 0.01) < 0.004


@pytest.mark.gpu
def test_input_map_prediction_and_gradient(gpu):
    from mcscram_amd.surrogate import Mlp
    n_in, h = 70, 64
    rng = np.random.default_rng(12)
    scale = (0.5 + rng.random(n_in)).astype(np.float32)
    shift = (-2.0 * rng.random(n_in)).astype(np.float32)
    cap = np.exp(rng.random(n_in) * 3).astype(np.float32)
    net = Mlp(n_in, h, seed=5)
    p = net.params()
    x = rng.standard_normal((128, n_in)).astype(np.float32)
    y = rng.standard_normal(128).astype(np.float32)
    net.set_input_map(scale, shift, cap)
    u = ref.apply_input_map(x, scale, shift, cap)
    assert np.allclose(net.predict(x), ref.predict(p, u, n_in, h), rtol=1e-2, atol=1e-3)
    loss, g = net.loss_grad(x, y)
    rloss, rg = ref.loss_and_grad(p, u, y, n_in, h)
    assert abs(loss - rloss) <= 1e-3 * max(1.0, rloss)
    assert np.linalg.norm(g - rg) <= 2e-2 * np.linalg.norm(rg)
    # Cleared map: the raw inputs again.
    net.set_input_map()
    assert np.allclose(net.predict(x), ref.predict(p, x, n_in, h), rtol=1e-2, atol=1e-3)
    with pytest.raises(Exception):
        net.set_input_map(scale, shift, None)


@pytest.mark.gpu
def test_ensemble_save_and_load_reproduce_predictions(gpu, tmp_path):
    from mcscram_amd.surrogate import Ensemble
    rng = np.random.default_rng(21)
    x = rng.standard_normal((512, 70)).astype(np.float32)
    y = x[:, :5].sum(axis=1)
    scale = (0.5 + rng.random(70)).astype(np.float32)
    shift = (-rng.random(70)).astype(np.float32)
    cap = np.full(70, 4.0, dtype=np.float32)
    ens = Ensemble(70, hidden=64, members=2, seed=4, input_map=(scale, shift, cap))
    ens.fit(x, y, epochs=3, batch=128)
    path = str(tmp_path / "ens.npz")
    ens.save(path, {"pairs": 512})
    back, info = Ensemble.load(path, input_map=(scale, shift, cap))
    assert info == {"pairs": 512}
    # Inference multiplies BF16 weights: saving their BF16 values loses nothing.
    assert np.array_equal(back.predict(x), ens.predict(x))


def test_bf16_bit_patterns_round_like_the_oracle():
    from mcscram_amd.surrogate import _from_bf16_bits, _to_bf16_bits
    x = np.random.default_rng(3).standard_normal(200_000).astype(np.float32) * 7.0
    x[:4] = [0.0, -0.0, 1.0, 3.0e38]
    assert np.array_equal(_from_bf16_bits(_to_bf16_bits(x)), ref.bf16(x).astype(np.float32))
