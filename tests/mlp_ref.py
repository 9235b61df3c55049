"""NumPy reference of the cost-model MLP (test infrastructure): the same
parameter layout and mean-squared-error gradient as the HIP kernels
(8 -> 32 -> 32 -> 1, ReLU, inputs standardised by stored mean / inv_std)."""
import numpy as np

F, H1, H2 = 8, 32, 32
NTRAIN = F * H1 + H1 + H1 * H2 + H2 + H2 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H1].reshape(F, H1); o += F * H1
    b1 = p[o:o + H1]; o += H1
    W2 = p[o:o + H1 * H2].reshape(H1, H2); o += H1 * H2
    b2 = p[o:o + H2]; o += H2
    W3 = p[o:o + H2]; o += H2
    b3 = p[o]; o += 1
    mean = p[o:o + F]; o += F
    istd = p[o:o + F]
    return W1, b1, W2, b2, W3, b3, mean, istd


def forward(p, X):
    W1, b1, W2, b2, W3, b3, mean, istd = unpack(p)
    xn = (np.asarray(X, dtype=np.float64) - mean) * istd
    z1 = xn @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ W3 + b3, (xn, h1, h2)


def loss(p, X, Y):
    y, _ = forward(p, X)
    return float(np.mean((y - np.asarray(Y, dtype=np.float64)) ** 2))


def gradient(p, X, Y):
    W1, b1, W2, b2, W3, b3, mean, istd = unpack(p)
    y, (xn, h1, h2) = forward(p, X)
    n = len(y)
    dy = 2.0 * (y - np.asarray(Y, dtype=np.float64)) / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, W3) * (h2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1 = xn.T @ d1
    gb1 = d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
