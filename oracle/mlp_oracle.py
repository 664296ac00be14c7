"""CPU oracle for the warm-start MLP kernels.

TEST INFRASTRUCTURE ONLY (imported by tests/ alone).  A numpy restatement of
the forward pass, the masked-MSE gradients and one Adam step (Kingma & Ba,
ICLR 2015, Algorithm 1, with decoupled weight decay), in float64 with BF16
rounding (round to nearest even) applied at the same points as the device:
inputs, weights, hidden activations and the back-propagated dZ that enter
the matrix products.  The device accumulates in FP32, so the two agree to
FP32 accumulation error, except where an FP32/FP64 value falls on the other
side of a BF16 rounding boundary.
"""
import numpy as np


def bf16(x):
    x32 = np.asarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(W, b, X):
    A = bf16(X)
    acts, Zs = [A], []
    for l in range(len(W)):
        Z = A @ bf16(W[l]) + b[l]
        Zs.append(Z)
        if l < len(W) - 1:
            A = bf16(silu(Z))
            acts.append(A)
    return Zs[-1], acts, Zs


def loss_and_grads(W, b, X, T, mask=None):
    Y, acts, Zs = forward(W, b, X)
    w = np.ones_like(Y) if mask is None else mask
    count = w.sum()
    d = Y - T
    loss = float((w * d * d).sum() / count)
    dZ = 2.0 * w * d / count
    gW, gb = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        gW[l] = acts[l].T @ bf16(dZ)
        gb[l] = dZ.sum(axis=0)
        if l == 0:
            break
        dA = bf16(dZ) @ bf16(W[l]).T
        dZ = dA * dsilu(Zs[l - 1])
    return loss, Y, gW, gb


def adam(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    g = g + wd * w
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
