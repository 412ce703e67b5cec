"""numpy restatement of the learned warm-start network (test oracle only).

The network has no counterpart in the reference: it maps a log-scaled
diffraction pattern (n x n) to the projected phase of a patch x patch region
around the probe and is trained on patterns that the multislice forward model
simulates.  This module is the checker for its GPU kernels.  Layers
(activations pixel-major, channels innermost):

    conv 4x4 / 2, 1 -> 16,  ReLU      n   -> n/2
    conv 4x4 / 2, 16 -> 32, ReLU      n/2 -> n/4
    conv 4x4 / 2, 32 -> 32, ReLU      n/4 -> n/8
    conv 4x4 / 2, 32 -> 64, ReLU      n/8 -> n/16
    dense (n/16)^2 * 64 -> patch^2, ReLU
    dense patch^2 -> patch^2 (the phase patch)

A convolution is im2col (column (ky*4 + kx)*C + c) followed by cols x W^T with
W (Cout, 16*Cin).  With ``bf16=True`` every GEMM operand is rounded to
bfloat16 (round to nearest even) exactly where the GPU rounds it: the im2col
columns, the BF16 activations fed to dense layers, the weights and the output
gradients; products accumulate in FP32.  With ``bf16=False`` and float64
inputs it is the exact network, used for the finite-difference checks that pin
the backward pass.  The loss is the mean squared error over all outputs.
"""
from __future__ import annotations

import numpy as np

KS, STRIDE, PAD = 4, 2, 1


def layer_spec(n_in, patch):
    """[(kind, cin, cout, relu, in_hw)] of the network for n x n inputs and patch x patch outputs."""
    assert n_in % 16 == 0
    spec, hw = [], n_in
    for cin, cout in ((1, 16), (16, 32), (32, 32), (32, 64)):
        spec.append(("conv", cin, cout, True, hw))
        hw //= 2
    spec.append(("dense", hw * hw * 64, patch * patch, True, 1))
    spec.append(("dense", patch * patch, patch * patch, False, 1))
    return spec


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, u & 0xFFFF0000, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def im2col(X, ks=KS, st=STRIDE, pad=PAD):
    """X (B, H, W, C) -> cols (B*Ho*Wo, ks*ks*C), column (ky*ks + kx)*C + c."""
    B, H, W, C = X.shape
    Ho, Wo = (H + 2 * pad - ks) // st + 1, (W + 2 * pad - ks) // st + 1
    Xp = np.zeros((B, H + 2 * pad, W + 2 * pad, C), X.dtype)
    Xp[:, pad:pad + H, pad:pad + W] = X
    cols = np.empty((B, Ho, Wo, ks, ks, C), X.dtype)
    for ky in range(ks):
        for kx in range(ks):
            cols[:, :, :, ky, kx] = Xp[:, ky:ky + st * Ho:st, kx:kx + st * Wo:st]
    return cols.reshape(B * Ho * Wo, ks * ks * C)


def col2im(dcols, shape, ks=KS, st=STRIDE, pad=PAD):
    """The adjoint of im2col: dX (B, H, W, C)."""
    B, H, W, C = shape
    Ho, Wo = (H + 2 * pad - ks) // st + 1, (W + 2 * pad - ks) // st + 1
    d = dcols.reshape(B, Ho, Wo, ks, ks, C)
    dXp = np.zeros((B, H + 2 * pad, W + 2 * pad, C), dcols.dtype)
    for ky in range(ks):
        for kx in range(ks):
            dXp[:, ky:ky + st * Ho:st, kx:kx + st * Wo:st] += d[:, :, :, ky, kx]
    return dXp[:, pad:pad + H, pad:pad + W]


def _r(x, bf16):
    return bf16_round(x) if bf16 else x


def forward(params, x, n_in, patch, bf16=True):
    """x (B, n, n) -> (out (B, patch^2), cache).  params: [(W (Cout, Kc), b (Cout,))]."""
    spec = layer_spec(n_in, patch)
    B = x.shape[0]
    dt = np.float32 if bf16 else np.float64
    A = x.reshape(B, n_in, n_in, 1).astype(dt)
    cache = []
    for (kind, cin, cout, relu, hw), (W, b) in zip(spec, params):
        if kind == "conv":
            cols = _r(im2col(A), bf16)
        else:
            cols = _r(A.reshape(B, -1), bf16)
        Y = (cols @ _r(W.astype(dt), bf16).T).astype(dt) + b.astype(dt)
        An = np.maximum(Y, 0) if relu else Y
        cache.append((kind, cols, An, A.shape))
        A = An.reshape(B, hw // 2, hw // 2, cout) if kind == "conv" else An
    return A.reshape(B, -1), cache


def loss_and_grads(params, x, target, n_in, patch, bf16=True):
    """Mean squared error and its gradients [(dW, db)] (the GPU's backward pass)."""
    spec = layer_spec(n_in, patch)
    out, cache = forward(params, x, n_in, patch, bf16)
    dt = out.dtype.type
    n = out.size
    diff = out - target.reshape(out.shape).astype(dt)
    loss = float(np.sum(diff.astype(np.float64) ** 2) / n)
    dA = (dt(2.0) / dt(n) * diff) if bf16 else 2.0 * diff / n
    grads = [None] * len(spec)
    for li in range(len(spec) - 1, -1, -1):
        kind, cin, cout, relu, hw = spec[li]
        _, cols, A, in_shape = cache[li]
        dY = (dA.reshape(A.shape) * (A > 0)) if relu else dA.reshape(A.shape)
        dY2 = dY.reshape(-1, cout)
        db = dY2.sum(axis=0)
        dYr = _r(dY2, bf16)
        dW = (dYr.T @ cols).astype(dt)
        grads[li] = (dW, db.astype(dt))
        if li > 0:
            W = params[li][0]
            dcols = (dYr @ _r(W.astype(dt), bf16)).astype(dt)
            dA = col2im(dcols, in_shape) if kind == "conv" else dcols
    return loss, grads


def adam_step(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015) in float32; returns (w, m, v)."""
    f = np.float32
    m = f(b1) * m + f(1 - b1) * g
    v = f(b2) * v + f(1 - b2) * g * g
    c1, c2 = f(1 - b1 ** step), f(1 - b2 ** step)
    w = w - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return w.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
