"""CPU oracle of the surrogate MLP: forward, loss and backward in numpy
FP64.  TEST INFRASTRUCTURE ONLY (checker of the BF16 MFMA kernels).

With ``bf16=True`` it rounds to BF16 (nearest even) exactly where the GPU
kernels do: network inputs, weights, hidden activations, the stored
pre-activations used by SiLU', and the back-propagated errors that feed
matrix products; accumulation stays FP64 (the GPU accumulates in FP32, so
comparisons carry a tolerance).  With ``bf16=False`` it is the plain FP64
network, which the finite-difference test pins.

Parameter layout (PARAM_COUNT = 10448 floats): W1[16][64] b1[64] W2[64][64]
b2[64] W3[64][64] b3[64] W4[64][16] b4[16]; inputs 13 (padded to 16),
outputs 7 (padded to 16).
"""
import numpy as np

SHAPES = [("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (64, 16)), ("b4", (16,))]
PARAM_COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, bf16_mode=True):
    q = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    n = x.shape[0]
    X = np.zeros((n, 16))
    X[:, :13] = x
    X = q(X)
    Ws = [q(P["W1"]), q(P["W2"]), q(P["W3"]), q(P["W4"])]
    bs = [P["b1"], P["b2"], P["b3"], P["b4"]]
    cache = {"X": X, "Ws": Ws}
    h = X
    for layer in range(3):
        z = h @ Ws[layer] + bs[layer]
        cache[f"z{layer + 1}"] = q(z)
        h = q(_silu(z))
        cache[f"h{layer + 1}"] = h
    y = h @ Ws[3] + bs[3]
    return y[:, :7], cache


def loss_and_grad(p, x, t, bf16_mode=True):
    q = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    n = x.shape[0]
    y, c = forward(p, x, bf16_mode)
    e = y - t
    loss = float(np.sum(e * e) / (n * 7))
    dY = np.zeros((n, 16))
    dY[:, :7] = 2.0 * e / (n * 7)
    g = {"b4": dY.sum(0)}
    dYq = q(dY)
    g["W4"] = c["h3"].T @ dYq
    d = dYq
    Wn = c["Ws"][3]
    for layer in (3, 2, 1):
        dH = d @ Wn.T
        dZ = dH * _silu_grad(c[f"z{layer}"])
        g[f"b{layer}"] = dZ.sum(0)
        dZq = q(dZ)
        inp = c["X"] if layer == 1 else c[f"h{layer - 1}"]
        g[f"W{layer}"] = inp.T @ dZq
        d = dZq
        Wn = c["Ws"][layer - 1]
    return loss, pack(g)
