"""CPU restatement (float64, numpy) of the format-selection MLP
24 -> 64 -> 64 -> 10 with ReLU: forward pass, weighted squared loss and
its analytic gradient, plus a central-difference gradient for pinning.

TEST INFRASTRUCTURE ONLY: tests/ compare the HIP training kernels'
gradients with this restatement.

Parameter layout (one flat array): W1[24][64] b1[64] W2[64][64] b2[64]
W3[64][10] b3[10], rows indexed by the input unit.
"""
import numpy as np

IN, HID, OUT = 24, 64, 10
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)),
          ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def forward(p, X):
    w = unpack(p)
    a1 = X @ w["W1"] + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ w["W3"] + w["b3"]
    return y, (X, a1, h1, a2, h2)


def loss(p, X, T, wout):
    y, _ = forward(p, X)
    return float(np.sum(np.asarray(wout) * (y - T) ** 2) / len(X))


def loss_grad(p, X, T, wout):
    """(loss, gradient) of (1/n) sum_s sum_o w_o (y_so - t_so)^2."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = len(X)
    w = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    e = y - T
    L = float(np.sum(np.asarray(wout) * e ** 2) / n)
    dy = 2.0 * np.asarray(wout) * e / n
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"].T) * (a2 > 0)
    g["W2"], g["b2"] = h1.T @ dh2, dh2.sum(0)
    dh1 = (dh2 @ w["W2"].T) * (a1 > 0)
    g["W1"], g["b1"] = X.T @ dh1, dh1.sum(0)
    return L, np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])


def numeric_grad(p, X, T, wout, idx, h=1e-6):
    """Central differences of the loss for the parameters in idx."""
    p = np.asarray(p, dtype=np.float64)
    out = np.zeros(len(idx))
    for k, i in enumerate(idx):
        q = p.copy()
        q[i] += h
        lp = loss(q, X, T, wout)
        q[i] -= 2 * h
        lm = loss(q, X, T, wout)
        out[k] = (lp - lm) / (2 * h)
    return out
