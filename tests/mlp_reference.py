"""FP64 numpy reference of the beat surrogate (test infrastructure): the same
network, loss and Adam step as include/ryr_mlp.h states them, written out
plainly so the GPU kernels can be checked against it.  Pinned by
finite differences in tests/test_mlp.py."""
import numpy as np

NIN = 11
C0, C1 = 0.7978845608028654, 0.044715


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[0:2048].reshape(32, 64)[:NIN], p[2048:2112], p[2112:6208].reshape(64, 64), p[6208:6272],
            p[6272:7296].reshape(64, 16)[:, :2], p[7296:7298])


def gelu(x):
    th = np.tanh(C0 * (x + C1 * x ** 3))
    return 0.5 * x * (1.0 + th), 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * C0 * (1.0 + 3.0 * C1 * x * x)


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = x @ W1 + b1
    a1, d1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2, d2 = gelu(z2)
    z3 = a2 @ W3 + b3
    return z3, (x, a1, d1, a2, d2)


def loss_and_grad(p, x, y, t, mask, s=None, w_apd=1.0):
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    s = np.ones(n) if s is None else np.asarray(s, dtype=np.float64)
    sw = s / s.sum()
    z3, (x, a1, d1, a2, d2) = forward(p, x)
    z, apd = z3[:, 0], z3[:, 1]
    bce = np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))
    e = apd - t
    loss = float(np.sum(sw * (bce + w_apd * mask * e * e)))
    dz3 = np.zeros((n, 16))
    dz3[:, 0] = sw * (1.0 / (1.0 + np.exp(-z)) - y)
    dz3[:, 1] = sw * 2.0 * w_apd * mask * e
    W1, b1, W2, b2, W3, b3 = unpack(p)
    g = np.zeros(7312)
    gW3 = g[6272:7296].reshape(64, 16)
    gW3[:] = a2.T @ dz3
    g[7296:7312] = dz3.sum(0)
    dz2 = (dz3[:, :2] @ W3.T) * d2
    g[2112:6208] = (a1.T @ dz2).ravel()
    g[6208:6272] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * d1
    gW1 = g[0:2048].reshape(32, 64)
    gW1[:NIN] = x.T @ dz1
    g[2048:2112] = dz1.sum(0)
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
