"""CPU oracle of the cost model's MLP (12 -> 64 -> 64 -> 2, ReLU) in float64.

TEST INFRASTRUCTURE ONLY.  Forward pass, loss and analytic gradient written directly from the
definitions (loss = mean over samples of 0.5 (y0 - t0)^2 + 0.5 w_exit (y1 - t1)^2); the analytic
gradient is itself checked against central finite differences in the tests.
Parameter layout: W1[64][12] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
"""
import numpy as np

NIN, NH, NOUT = 12, 64, 2
SIZES = [(NH, NIN), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SIZES)


def unpack(w):
    out, o = [], 0
    for s in SIZES:
        k = int(np.prod(s))
        out.append(np.asarray(w[o:o + k], np.float64).reshape(s))
        o += k
    return out


def pack(parts):
    return np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])


def forward(w, z):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    a1 = z @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3.T + b3, (a1, h1, a2, h2)


def loss_grad(w, z, t, w_exit):
    z = np.asarray(z, np.float64)
    t = np.asarray(t, np.float64)
    n = len(z)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    y, (a1, h1, a2, h2) = forward(w, z)
    e = y - t
    wt = np.array([1.0, w_exit])
    loss = float(np.mean(0.5 * (e * e) @ wt))
    g3 = e * wt / n
    gW3 = g3.T @ h2
    gb3 = g3.sum(0)
    g2 = (g3 @ W3) * (a2 > 0)
    gW2 = g2.T @ h1
    gb2 = g2.sum(0)
    g1 = (g2 @ W2) * (a1 > 0)
    gW1 = g1.T @ z
    gb1 = g1.sum(0)
    return loss, pack([gW1, gb1, gW2, gb2, gW3, gb3])


def loss_only(w, z, t, w_exit):
    y, _ = forward(w, np.asarray(z, np.float64))
    e = y - np.asarray(t, np.float64)
    return float(np.mean(0.5 * (e[:, 0] ** 2 + w_exit * e[:, 1] ** 2)))
