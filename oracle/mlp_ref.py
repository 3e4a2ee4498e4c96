"""CPU restatement of the dt0 / method-class / cost MLP (test infrastructure only).

The network has no counterpart in the reference; it is judged by the
solver it feeds (the initdt heuristic and AutoTsit5 switching remain the
fallback).  This file is the checker for the HIP forward/backward kernels:
the same arithmetic in float64 numpy, with the BF16 rounding applied at the
points where the kernels round (normalised inputs, weights, hidden
activations, and the pre-activation gradients that feed matrix cores).

Network: 16 inputs -> 64 -> 64 -> 64 -> 4, GELU (tanh form) after each
hidden layer.  Parameter vector layout (float32): W1[16x64], b1[64],
W2[64x64], b2[64], W3[64x64], b3[64], W4[64x4], b4[4]; W[in][out] row-major.
Outputs: log10 dt0, stiffness logit, log2 cost, never-stiff logit.
Loss: w_dt * mean((y0 - target)^2) + w_cls * mean(BCE-with-logits(y1, label))
      + w_cost * mean((y2 - cost)^2) + w_never * mean(BCE-with-logits(y3, never))
      (the cost and never terms dropped when cost / never is None).
"""
import numpy as np

NIN, NHID, NOUT = 16, 64, 4
SHAPES = [("W1", (NIN, NHID)), ("b1", (NHID,)), ("W2", (NHID, NHID)), ("b2", (NHID,)),
          ("W3", (NHID, NHID)), ("b3", (NHID,)), ("W4", (NHID, NOUT)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
GELU_C = 0.7978845608028654  # sqrt(2/pi)
GELU_A = 0.044715


def bf16(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    out, k = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[k:k + n], dtype=np.float64).reshape(shp)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def gelu(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def _ident(x):
    return np.asarray(x, dtype=np.float64)


def normalize(X, mean, std, rnd=bf16):
    rstd = (1.0 / np.asarray(std, dtype=np.float32)).astype(np.float32)
    xn = (np.asarray(X, dtype=np.float32) - np.asarray(mean, dtype=np.float32)) * rstd
    return rnd(xn)


def forward(theta, X, mean, std, exact=False):
    """exact=True: no BF16 rounding anywhere (float64 throughout)."""
    rnd = _ident if exact else bf16
    P = unpack(theta)
    W = {k: rnd(P[k]) if k.startswith("W") else P[k] for k in P}
    h = normalize(X, mean, std, rnd)
    cache = {"H0": h, "rnd": rnd}
    for l in (1, 2, 3):
        z = h @ W[f"W{l}"] + P[f"b{l}"]
        h = rnd(gelu(z))
        cache[f"Z{l}"] = z
        cache[f"H{l}"] = h
    y = h @ W["W4"] + P["b4"]
    return y, cache, W


def _bce(y, c):
    return np.maximum(y, 0) - y * c + np.log1p(np.exp(-np.abs(y)))


def loss(theta, X, target, label, mean, std, w_dt=1.0, w_cls=1.0, exact=False, cost=None,
         w_cost=1.0, never=None, w_never=1.0):
    y, _, _ = forward(theta, X, mean, std, exact)
    e = y[:, 0] - target
    bce = _bce(y[:, 1], label)
    ec = np.zeros(len(X)) if cost is None else y[:, 2] - cost
    wc = 0.0 if cost is None else w_cost
    bn = np.zeros(len(X)) if never is None else _bce(y[:, 3], never)
    wn = 0.0 if never is None else w_never
    return (w_dt * np.mean(e * e) + w_cls * np.mean(bce) + wc * np.mean(ec * ec)
            + wn * np.mean(bn))


def loss_and_grad(theta, X, target, label, mean, std, w_dt=1.0, w_cls=1.0, exact=False,
                  cost=None, w_cost=1.0, never=None, w_never=1.0):
    y, c, W = forward(theta, X, mean, std, exact)
    bf16_ = c["rnd"]
    B = len(X)
    e = y[:, 0] - target
    sig = 1.0 / (1.0 + np.exp(-y[:, 1]))
    bce = _bce(y[:, 1], label)
    ec = np.zeros(B) if cost is None else y[:, 2] - cost
    wc = 0.0 if cost is None else w_cost
    nv = np.zeros(B) if never is None else np.asarray(never, dtype=np.float64)
    wn = 0.0 if never is None else w_never
    sig3 = 1.0 / (1.0 + np.exp(-y[:, 3]))
    bn = _bce(y[:, 3], nv)
    loss = (w_dt * np.mean(e * e) + w_cls * np.mean(bce) + wc * np.mean(ec * ec)
            + wn * np.mean(bn))
    dz4 = np.stack([2.0 * e * w_dt / B, (sig - label) * w_cls / B, 2.0 * ec * wc / B,
                    (sig3 - nv) * wn / B], axis=1)
    g = {"W4": c["H3"].T @ bf16_(dz4), "b4": dz4.sum(0)}
    dz_b = bf16_(dz4)
    for l in (3, 2, 1):
        dh = dz_b @ W[f"W{l + 1}"].T
        dz = dh * gelu_grad(c[f"Z{l}"])
        g[f"b{l}"] = dz.sum(0)
        dz_b = bf16_(dz)
        g[f"W{l}"] = c[f"H{l - 1}"].T @ dz_b
    return loss, pack(g), y


def init_params(seed=0):
    """Xavier-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            d[name] = rng.uniform(-lim, lim, shp)
        else:
            d[name] = np.zeros(shp)
    return pack(d).astype(np.float32)
