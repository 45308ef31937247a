"""GPU checks of the selector MLP's HIP kernels (forward, backward, Adam,
training loop) against the CPU oracle oracle/mlp_oracle.py, at the selector's
shape 64-128-128-25 with BF16 weights and FP32 accumulation."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu
SIZES = (64, 128, 128, 25)


def _data(B, seed):
    r = np.random.default_rng(seed)
    X = r.normal(0, 1, (B, SIZES[0])).astype(np.float32)
    T = r.normal(0, 1, (B, SIZES[-1])).astype(np.float32)
    M = (r.random((B, SIZES[-1])) < 0.9).astype(np.float32)
    return X, T, M


def test_params_roundtrip_and_layout(sp, ctx):
    m = sp.Mlp(ctx, SIZES, seed=1)
    assert m.nparams == mo.layout(SIZES)[2] == 28057
    p = m.get_params()
    assert np.all(np.isfinite(p))
    q = np.random.default_rng(0).normal(0, 0.1, m.nparams).astype(np.float32)
    m.set_params(q)
    np.testing.assert_array_equal(m.get_params(), q)


@pytest.mark.parametrize("B", [1, 37, 256])
def test_forward_matches_oracle(sp, ctx, B):
    m = sp.Mlp(ctx, SIZES, seed=2)
    X, _, _ = _data(B, 10 + B)
    y = m.predict(X)
    want, _ = mo.forward(m.get_params(), SIZES, X)
    scale = np.abs(want).max()
    # FP32 accumulation over <= 128 terms vs float64 on identical BF16 weights
    np.testing.assert_allclose(y, want, rtol=0, atol=2e-5 * scale)


@pytest.mark.parametrize("B", [16, 200])
def test_gradient_matches_oracle(sp, ctx, B):
    m = sp.Mlp(ctx, SIZES, seed=3)
    X, T, M = _data(B, 20 + B)
    L, g = m.loss_grad(X, T, M)
    Lw, gw = mo.loss_grad(m.get_params(), SIZES, X, T, M)
    assert abs(L - Lw) <= 1e-5 * abs(Lw)
    _, offs, n = mo.layout(SIZES)
    bounds = offs + [n]
    for j in range(6):
        a, b = bounds[j], bounds[j + 1]
        err = np.abs(g[a:b] - gw[a:b]).max()
        assert err <= 1e-4 * np.abs(gw[a:b]).max() + 1e-9, (j, err)


def test_adam_step_matches_oracle(sp, ctx):
    m = sp.Mlp(ctx, SIZES, seed=4)
    X, T, M = _data(64, 5)
    p0 = m.get_params().astype(np.float64)
    _, g = m.loss_grad(X, T, M)
    m.adam(lr=1e-3)
    p1 = m.get_params()
    want, mm, vv = mo.adam(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, lr=1e-3)
    np.testing.assert_allclose(p1, want, rtol=0, atol=1e-6)
    # second step continues the moments
    _, g2 = m.loss_grad(X, T, M)
    m.adam(lr=1e-3)
    want2, _, _ = mo.adam(want, g2.astype(np.float64), mm, vv, 2, lr=1e-3)
    np.testing.assert_allclose(m.get_params(), want2, rtol=0, atol=2e-6)


def test_training_reduces_loss(sp, ctx):
    r = np.random.default_rng(6)
    X = r.normal(0, 1, (512, SIZES[0])).astype(np.float32)
    Wt = r.normal(0, 0.3, (SIZES[0], SIZES[-1]))
    T = np.tanh(X @ Wt).astype(np.float32)
    m = sp.Mlp(ctx, SIZES, seed=7)
    L0, _ = m.loss_grad(X, T)
    Lf = m.train(X, T, epochs=60, batch=64, lr=2e-3, seed=1)
    L1, _ = m.loss_grad(X, T)
    assert Lf < 0.2 * L0 and L1 < 0.2 * L0, (L0, Lf, L1)
    # identical seeds give identical training (fixed reduction order)
    m2 = sp.Mlp(ctx, SIZES, seed=7)
    m2.train(X, T, epochs=60, batch=64, lr=2e-3, seed=1)
    np.testing.assert_array_equal(m.get_params(), m2.get_params())
