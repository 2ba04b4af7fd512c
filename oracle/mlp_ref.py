"""ORACLE — test infrastructure only.  numpy restatement of the warm-start
network (12 -> 128 -> 128 -> 128 -> 3, GELU) used to check the device
kernels.  The network has no counterpart in the reference, so this is the
specification: the forward pass, the mean-squared-error loss and its exact
backward pass, in float64, optionally with the device's BF16 rounding points
(weights, layer inputs, GELU outputs and derivatives, and the backward
operands are rounded to BF16 where the device feeds its matrix cores).  The
exact (unrounded) backward pass is pinned against central finite differences
in tests/test_warm_start.py.
"""
import numpy as np
from scipy.special import erf

IN, H, OUT = 12, 128, 3
SHAPES = [("W0", (IN, H)), ("b0", (H,)), ("W1", (H, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H, OUT)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAM == 35075


def bf16(a):
    """Round-to-nearest-even to BF16, returned as float64 values."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES])


def init_params(seed: int = 0, out_scale: float = 0.1):
    """He-normal weights, zero biases; the output layer scaled by out_scale."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), size=shape)
        else:
            d[name] = np.zeros(shape)
    d["W3"] *= out_scale
    return pack(d).astype(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _r(a, rounding):
    return bf16(a) if rounding else np.asarray(a, np.float64)


def forward(p, X, rounding=True):
    """Predictions [n, 3] and the cached layer values."""
    w = unpack(p)
    a = [_r(X, rounding)]
    g = []
    for l in range(3):
        z = a[-1] @ _r(w[f"W{l}"], rounding) + w[f"b{l}"]
        a.append(_r(gelu(z), rounding))
        g.append(_r(gelu_d(z), rounding))
    y = a[3] @ _r(w["W3"], rounding) + w["b3"]
    return y, a, g


def loss_and_grad(p, X, T, rounding=True):
    """Mean squared error over the 3 outputs of all samples and its gradient
    (master parameter order)."""
    w = unpack(p)
    y, a, g = forward(p, X, rounding)
    n = len(X)
    e = y - T
    loss = float((e * e).sum() / (3 * n))
    dz = 2.0 * e / (3 * n)
    grads = {"b3": dz.sum(0)}
    dzr = _r(dz, rounding)
    grads["W3"] = a[3].T @ dzr
    for l in (3, 2, 1):
        da = dzr @ _r(w[f"W{l}"], rounding).T
        dz = da * g[l - 1]
        grads[f"b{l - 1}"] = dz.sum(0)
        dzr = _r(dz, rounding)
        grads[f"W{l - 1}"] = a[l - 1].T @ dzr
    return loss, pack(grads)
