"""CPU restatement (numpy, float64) of the warm-start network for checking the GPU kernels.

Same architecture and parameter layout as sb_mlp (64 -> 256 -> 256 -> 256 -> 256 -> 1,
GELU in the sigmoid form x s(1.702 x) of Hendrycks & Gimpel, s the logistic function),
and the same rounding points: the GEMM operands are rounded to bf16
(inputs, weights, activations, back-propagated gradients) and everything else is
computed here in float64.  Test infrastructure only.
"""
import numpy as np

LAYERS = [(64, 256), (256, 256), (256, 256), (256, 256), (256, 1)]  # (in, out)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); returns float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = [], 0
    for fi, fo in LAYERS:
        W = p[o:o + fo * fi].reshape(fo, fi).astype(np.float64)
        o += fo * fi
        b = p[o:o + fo].astype(np.float64)
        o += fo
        out.append((W, b))
    return out


def _s(z):
    return 1.0 / (1.0 + np.exp(-1.702 * z))


def gelu(z):
    return z * _s(z)


def gelu_grad(z):
    s = _s(z)
    return s + 1.702 * z * s * (1.0 - s)


def forward(p, X):
    layers = unpack(p)
    H = bf16(X)
    Zs, Hs = [], [H]
    for W, b in layers[:4]:
        Z = H @ bf16(W).T + b
        H = bf16(gelu(Z))
        Zs.append(Z)
        Hs.append(H)
    W4, b4 = layers[4]
    y = H @ bf16(W4[0]) + b4[0]
    return y, Zs, Hs


def gradients(p, X, t, transpose_first_weight=False):
    """Mean squared error and its gradient in the parameter layout.  With
    transpose_first_weight the restatement carries a planted mistake (used to show
    that the comparison detects one)."""
    if transpose_first_weight:
        p = p.copy()
        W0 = p[:256 * 64].reshape(256, 64)
        p[:256 * 64] = np.ascontiguousarray(W0.reshape(64, 256).T).reshape(-1)
    layers = unpack(p)
    n = X.shape[0]
    y, Zs, Hs = forward(p, X)
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    grads = [None] * 5
    W4, _ = layers[4]
    grads[4] = (dy @ Hs[4], np.array([dy.sum()]))
    dZ = bf16(np.outer(dy, bf16(W4[0])) * gelu_grad(Zs[3]))
    for l in range(3, -1, -1):
        grads[l] = (dZ.T @ Hs[l], dZ.sum(axis=0))
        if l > 0:
            dH = dZ @ bf16(layers[l][0])
            dZ = bf16(dH * gelu_grad(Zs[l - 1]))
    flat = np.concatenate([np.concatenate([g[0].reshape(-1), g[1].reshape(-1)]) for g in grads])
    return loss, flat, y


def layer_slices():
    out, o = [], 0
    for fi, fo in LAYERS:
        out.append((slice(o, o + fo * fi), slice(o + fo * fi, o + fo * fi + fo)))
        o += fo * fi + fo
    return out
