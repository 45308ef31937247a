"""CPU pins of the selector-MLP oracle (oracle/mlp_oracle.py): the analytic
gradient against central finite differences (and a planted mistake that the
check must catch), BF16 nearest-even rounding against hand-worked cases and
torch's CPU cast, and Adam's first step against its closed form."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

SIZES = (6, 9, 7, 5)


def _case(seed=3, B=11):
    r = np.random.default_rng(seed)
    _, _, n = mo.layout(SIZES)
    p = r.normal(0, 0.6, n)
    X = r.normal(0, 1, (B, SIZES[0]))
    T = r.normal(0, 1, (B, SIZES[-1]))
    M = (r.random((B, SIZES[-1])) < 0.8).astype(np.float64) * r.uniform(0.5, 2.0, (B, SIZES[-1]))
    return p, X, T, M


def _fd(p, X, T, M, h=1e-6):
    g = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        g[i] = (mo.loss(p + e, SIZES, X, T, M, rounded=False) - mo.loss(p - e, SIZES, X, T, M, rounded=False)) / (2 * h)
    return g


def test_gradient_matches_finite_differences():
    p, X, T, M = _case()
    L, g = mo.loss_grad(p, SIZES, X, T, M, rounded=False)
    assert abs(L - mo.loss(p, SIZES, X, T, M, rounded=False)) < 1e-12
    fd = _fd(p, X, T, M)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_fd_check_catches_dropped_relu_mask():
    # planted mistake: backward without relu'(z2); the FD check must reject it
    p, X, T, M = _case(seed=4)
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, SIZES, rounded=False)
    y, (Xf, z1, a1, z2, a2) = mo.forward(p, SIZES, X, rounded=False)
    d3 = 2.0 * M * M * (y - T) / y.size
    d2 = d3 @ W3  # mask missing
    gW2 = d2.T @ a1
    _, offs, _ = mo.layout(SIZES)
    fd = _fd(p, X, T, M)[offs[2]:offs[3]]
    assert not np.allclose(gW2.ravel(), fd, rtol=1e-5, atol=1e-8)


def test_mask_scales_gradient_squared():
    # with a constant mask c the loss and gradient scale by c^2
    p, X, T, _ = _case(seed=5)
    L1, g1 = mo.loss_grad(p, SIZES, X, T, None, rounded=False)
    L3, g3 = mo.loss_grad(p, SIZES, X, T, np.full(T.shape, 3.0), rounded=False)
    assert abs(L3 - 9 * L1) < 1e-10 * max(1, L3)
    np.testing.assert_allclose(g3, 9 * g1, rtol=1e-12, atol=1e-14)


def test_bf16_round_hand_cases():
    f = np.float32
    cases = [
        (1.0, 1.0),
        (1 + 2 ** -8, 1.0),                      # tie, even stays
        (1 + 3 * 2 ** -8, 1 + 2 ** -6),          # tie, odd rounds up
        (1 + 2 ** -8 + 2 ** -20, 1 + 2 ** -7),   # above the tie
        (-1.5, -1.5),
        (np.finfo(np.float32).max, np.inf),      # rounds past the largest BF16
        (2 ** -130, 2 ** -130),                  # subnormal with few bits
    ]
    got = mo.bf16_round(np.array([c[0] for c in cases], f))
    np.testing.assert_array_equal(got, np.array([c[1] for c in cases], f))
    assert np.isnan(mo.bf16_round(np.array([np.nan], f)))[0]


def test_bf16_round_matches_torch_cpu_cast():
    torch = pytest.importorskip("torch")
    r = np.random.default_rng(9)
    a = r.normal(0, 10, 100000).astype(np.float32)
    bits = a.view(np.uint32)
    bits[:2000] = (bits[:2000] & 0xFFFF0000) | 0x8000  # exact ties
    a = bits.view(np.float32)
    want = torch.from_numpy(a.copy()).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mo.bf16_round(a), want)


def test_adam_first_step_closed_form():
    r = np.random.default_rng(2)
    p = r.normal(0, 1, 50)
    g = r.normal(0, 1, 50)
    g[3] = 0.0
    lr = 0.01
    p1, m1, v1 = mo.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=lr, eps=0.0 + 1e-12)
    step = p - p1
    nz = g != 0
    np.testing.assert_allclose(step[nz], lr * np.sign(g[nz]), rtol=1e-9)
    assert step[3] == 0.0
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(v1, 0.001 * g * g)
