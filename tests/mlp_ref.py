"""FP64 numpy restatement of the rank-predictor network (test infrastructure).

Same parameter layout as the HIP kernels: row-major W[in][out] blocks
W1 18x64 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 64 | b4 1, tanh-GELU
hidden activations, scalar output, mean-squared-error loss.
"""
import numpy as np

NIN, NH = 18, 64
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NH,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
C = np.sqrt(2.0 / np.pi)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = p[o:o + k].reshape(shp)
        o += k
    return out


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C * (1 + 3 * 0.044715 * x * x)


def init(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    w = unpack(p)
    w["W1"][:] = rng.standard_normal((NIN, NH)) / np.sqrt(NIN)
    w["W2"][:] = rng.standard_normal((NH, NH)) / np.sqrt(NH)
    w["W3"][:] = rng.standard_normal((NH, NH)) / np.sqrt(NH)
    w["W4"][:] = rng.standard_normal(NH) / np.sqrt(NH)
    for b in ("b1", "b2", "b3"):
        w[b][:] = 0.1 * rng.standard_normal(NH)
    return p


def forward(p, X):
    w = unpack(p)
    z1 = X @ w["W1"] + w["b1"]
    a1 = gelu(z1)
    z2 = a1 @ w["W2"] + w["b2"]
    a2 = gelu(z2)
    z3 = a2 @ w["W3"] + w["b3"]
    a3 = gelu(z3)
    out = a3 @ w["W4"] + w["b4"][0]
    return out, (z1, a1, z2, a2, z3, a3)


def loss_grad(p, X, Y):
    w = unpack(p)
    out, (z1, a1, z2, a2, z3, a3) = forward(p, X)
    ns = X.shape[0]
    diff = out - Y
    loss = float(np.mean(diff ** 2))
    g = np.zeros_like(p)
    gw = unpack(g)
    dout = 2.0 * diff / ns
    gw["W4"][:] = a3.T @ dout
    gw["b4"][0] = dout.sum()
    d3 = np.outer(dout, w["W4"]) * gelu_grad(z3)
    gw["W3"][:] = a2.T @ d3
    gw["b3"][:] = d3.sum(0)
    d2 = (d3 @ w["W3"].T) * gelu_grad(z2)
    gw["W2"][:] = a1.T @ d2
    gw["b2"][:] = d2.sum(0)
    d1 = (d2 @ w["W2"].T) * gelu_grad(z1)
    gw["W1"][:] = X.T @ d1
    gw["b1"][:] = d1.sum(0)
    return loss, g


def adam(p, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
