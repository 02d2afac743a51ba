"""CPU restatement of the surrogate network's arithmetic -- test
infrastructure only (the GPU kernels in se2w_mlp.hip are the product; this
file is the checker of their forward pass, gradients and Adam step).

The network: linear layers with GELU (erf form) after every hidden layer,
mean-squared-error loss.  With bf16=True, values are rounded to bfloat16
(round to nearest even) at the points where the GPU stores BF16: the input,
the weights, every hidden activation and every pre-activation gradient.
Sums are formed in float64 (the GPU accumulates in float32).

Parameter layout (as the C ABI): for each layer W [out][in] row-major, then
b [out].
"""
import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def unpack(params, widths):
    out, off = [], 0
    for i, o in zip(widths[:-1], widths[1:]):
        W = np.asarray(params[off:off + o * i], dtype=np.float64).reshape(o, i)
        off += o * i
        b = np.asarray(params[off:off + o], dtype=np.float64)
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def forward(params, widths, x, bf=True):
    """Returns the output and the per-layer (input activation, pre-activation)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, widths)
    h = r(x)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        cache.append((h, z))
        h = r(gelu(z)) if l < len(layers) - 1 else z
    return h, cache


def loss_and_gradient(params, widths, x, target, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, widths)
    y, cache = forward(params, widths, x, bf)
    d = y - target
    loss = float(np.mean(d * d))
    dz = r(2.0 * d / d.size)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        h, _ = cache[l]
        W, _ = layers[l]
        grads[l] = (dz.T @ h, dz.sum(axis=0))
        if l > 0:
            dh = dz @ r(W)
            dz = r(dh * gelu_grad(cache[l - 1][1]))
    return loss, pack(grads)


def adam_step(params, m, v, grad, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (Kingma & Ba 2015, Algorithm 1), with decoupled weight
    decay wd (AdamW, Loshchilov & Hutter 2019); returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * (mh / (np.sqrt(vh) + eps) + wd * params), m, v


def init_params(widths, seed):
    """The C ABI's initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a
    splitmix64 stream (weights row by row, then the biases, layer by layer)."""
    state = [seed & 0xFFFFFFFFFFFFFFFF]
    mask = 0xFFFFFFFFFFFFFFFF

    def u01():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & mask
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        return (z >> 11) * (1.0 / 9007199254740992.0)

    layers = []
    for i, o in zip(widths[:-1], widths[1:]):
        bound = 1.0 / np.sqrt(i)
        W = np.array([[(2.0 * u01() - 1.0) * bound for _ in range(i)] for _ in range(o)], dtype=np.float32)
        b = np.array([(2.0 * u01() - 1.0) * bound for _ in range(o)], dtype=np.float32)
        layers.append((W.astype(np.float64), b.astype(np.float64)))
    return pack(layers)
