"""TEST INFRASTRUCTURE ONLY -- the CPU oracle of the variant-selector MLP.

Only tests/ and bench.py's CPU leg may import this module, and only to check
the library's HIP kernels (sp_mlp_* in libsparsep.so); the product never
calls it.  The MLP has no counterpart in the reference.

It restates, in float64 numpy, what the HIP kernels compute:
  * BF16 weights: each FP32 master weight rounded to nearest-even BF16
    (bf16_round), biases kept in FP32;
  * forward  z1 = x W1^T + b1, a1 = relu(z1), z2 = a1 W2^T + b2,
    a2 = relu(z2), y = a2 W3^T + b3;
  * loss     L = mean over the B*out outputs of (M (y - t))^2;
  * backward dY = 2 M^2 (y - t) / (B out), the gradient of a weight is the
    gradient of the BF16 copy used forward (straight-through), relu'(z) = [z > 0];
  * Adam     m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
             p -= lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).
Parameter layout: W1[h1][in] b1[h1] W2[h2][h1] b2[h2] W3[out][h2] b3[out].

Pinned by tests/test_mlp_oracle.py: the analytic gradient against central
finite differences with rounding off, bf16_round against hand-worked
nearest-even cases and torch's CPU bfloat16 cast, and Adam's first step
against its closed form (|update| = lr for every parameter with g != 0).
"""
import numpy as np


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even); returns float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x00400000) & 0xFFFF0000, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(np.shape(a))


def layout(sizes):
    i, h1, h2, o = sizes
    shapes = [(h1, i), (h1,), (h2, h1), (h2,), (o, h2), (o,)]
    offs, k = [], 0
    for s in shapes:
        offs.append(k)
        k += int(np.prod(s))
    return shapes, offs, k


def unpack(p, sizes, rounded=True):
    shapes, offs, _ = layout(sizes)
    out = []
    for j, (s, o) in enumerate(zip(shapes, offs)):
        v = np.asarray(p[o:o + int(np.prod(s))]).reshape(s)
        if rounded:
            v = v.astype(np.float32)
            if j % 2 == 0:
                v = bf16_round(v)
        out.append(v.astype(np.float64))
    return out


def forward(p, sizes, X, rounded=True):
    W1, b1, W2, b2, W3, b3 = unpack(p, sizes, rounded)
    X = np.asarray(X, np.float64)
    z1 = X @ W1.T + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0)
    y = a2 @ W3.T + b3
    return y, (X, z1, a1, z2, a2)


def loss(p, sizes, X, T, M=None, rounded=True):
    y, _ = forward(p, sizes, X, rounded)
    M = np.ones_like(y) if M is None else np.asarray(M, np.float64)
    return float(np.mean((M * (y - np.asarray(T, np.float64))) ** 2))


def loss_grad(p, sizes, X, T, M=None, rounded=True):
    """(loss, gradient over the flat parameter vector), float64."""
    W1, b1, W2, b2, W3, b3 = unpack(p, sizes, rounded)
    y, (X, z1, a1, z2, a2) = forward(p, sizes, X, rounded)
    M = np.ones_like(y) if M is None else np.asarray(M, np.float64)
    d = M * (y - np.asarray(T, np.float64))
    L = float(np.mean(d ** 2))
    d3 = 2.0 * M * d / y.size
    gW3, gb3 = d3.T @ a2, d3.sum(0)
    d2 = (d3 @ W3) * (z2 > 0)
    gW2, gb2 = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1, gb1 = d1.T @ X, d1.sum(0)
    g = np.concatenate([v.ravel() for v in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return L, g


def adam(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step at step number t (1-based); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
