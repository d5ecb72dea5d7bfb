"""FP64 reference of the warm-start network's forward pass, MSE loss and gradients.

TEST INFRASTRUCTURE ONLY: checks the GPU network kernels (chm_amd/csrc/mlp.hip).
The network is 16 -> 128 -> 128 -> 128 -> 11 with ReLU, normalised inputs and
targets, MSE over the 11 outputs.  With ``bf16=True`` the operands are rounded
to bfloat16 (round to nearest even) at the same points as the GPU kernels --
the normalised inputs, the weights, the hidden activations and the
back-propagated pre-activation gradients -- while the arithmetic stays FP64,
so what remains between the two is FP32 accumulation order.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT, OUTP = 16, 128, 11, 16
SHAPES = [(H, IN), (H, H), (H, H), (OUTP, H)]


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    """Flat parameter vector -> [(W, b)] per layer."""
    out, o = [], 0
    for (r, c) in SHAPES:
        W = params[o:o + r * c].reshape(r, c).astype(np.float64)
        o += r * c
        b = params[o:o + r].astype(np.float64)
        o += r
        out.append((W, b))
    assert o == params.size
    return out


def param_count() -> int:
    return sum(r * c + r for r, c in SHAPES)


def loss_and_grad(params, X, Y, norm, bf16: bool = True):
    """MSE loss (mean over samples and the 11 outputs, normalised targets) and the
    flat gradient, in FP64.  norm = (x_mean, x_std, y_mean, y_std)."""
    xm, xs, ym, ys = (np.asarray(v, dtype=np.float64) for v in norm)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params))
    Ws = [rnd(W) for W, _ in layers]
    bs = [b for _, b in layers]
    n = X.shape[0]
    a = rnd((np.asarray(X, np.float64) - xm) / xs)
    acts, zs = [a], []
    for l in range(4):
        z = acts[-1] @ Ws[l].T + bs[l]
        zs.append(z)
        if l < 3:
            acts.append(rnd(np.maximum(z, 0.0)))
    T = np.zeros((n, OUTP))
    T[:, :OUT] = (np.asarray(Y, np.float64) - ym) / ys
    d = zs[3] - T
    d[:, OUT:] = 0.0
    loss = float((d[:, :OUT] ** 2).sum() / (n * OUT))
    dz = rnd(2.0 * d / (n * OUT))
    grads = [None] * 4
    for l in range(3, -1, -1):
        gW = dz.T @ acts[l]
        gb = dz.sum(0)
        grads[l] = (gW, gb)
        if l > 0:
            da = dz @ Ws[l]
            dz = rnd(da * (zs[l - 1] > 0))
    flat = np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])
    return loss, flat


def predict(params, X, norm, bf16: bool = True):
    xm, xs, ym, ys = (np.asarray(v, dtype=np.float64) for v in norm)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params))
    a = rnd((np.asarray(X, np.float64) - xm) / xs)
    for l, (W, b) in enumerate(layers):
        z = a @ rnd(W).T + b
        a = rnd(np.maximum(z, 0.0)) if l < 3 else z
    return a[:, :OUT] * ys + ym
