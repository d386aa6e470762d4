"""float64 numpy restatement of the warm-start network (test infrastructure).

TEST INFRASTRUCTURE ONLY: the checker for the BF16 MFMA kernels of
muphfasa_amd/csrc/mfa_surrogate.hip (no reference counterpart exists; the
reference has no learned component).  Forward and backward of 3x3
convolutions (padding 1, stride 1/2, NHWC), ReLU, nearest upsampling with
additive skips and the masked MSE, written directly from their definitions;
its own gradients are pinned by central finite differences in
tests/test_surrogate_ref.py.
"""
from __future__ import annotations

import numpy as np


def conv3x3(x, w, b, stride=1):
    """x (N,H,W,Cin), w (Cout,3,3,Cin), b (Cout,) -> (N,Ho,Wo,Cout)."""
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    Ho, Wo = (H + stride - 1) // stride, (W + stride - 1) // stride
    xp = np.zeros((N, H + 2, W + 2, Cin))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((N, Ho, Wo, Cout))
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, ky:ky + stride * Ho:stride, kx:kx + stride * Wo:stride, :]
            y += np.einsum("nhwc,oc->nhwo", patch, w[:, ky, kx, :])
    return y + b


def conv3x3_bwd(x, w, dy, stride=1):
    """Gradients (dx, dw, db) of conv3x3 for upstream dy."""
    N, H, W, Cin = x.shape
    Ho, Wo = dy.shape[1], dy.shape[2]
    xp = np.zeros((N, H + 2, W + 2, Cin))
    xp[:, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            sl = (slice(None), slice(ky, ky + stride * Ho, stride), slice(kx, kx + stride * Wo, stride))
            dw[:, ky, kx, :] = np.einsum("nhwo,nhwc->oc", dy, xp[sl])
            dxp[sl] += np.einsum("nhwo,oc->nhwc", dy, w[:, ky, kx, :])
    return dxp[:, 1:-1, 1:-1], dw, dy.sum(axis=(0, 1, 2))


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def upsample_bwd(dy):
    N, H2, W2, C = dy.shape
    return dy.reshape(N, H2 // 2, 2, W2 // 2, 2, C).sum(axis=(2, 4))


def forward(params, layers, skips, x0):
    """params[name] = (w (Cout,3,3,Cin), b); returns (out, activations, pre)."""
    act = {"x0": x0}
    pre = {}
    for name, cin, cout, stride, relu, src in layers:
        if src in skips and src not in act:
            a, b = skips[src]
            act[src] = upsample(act[a]) + act[b]
        w, b = params[name]
        z = conv3x3(act[src], w, b, stride)
        pre[name] = z
        act[name] = np.maximum(z, 0.0) if relu else z
    return act["out"], act, pre


def mse(out, target, solid, nt):
    fluid = (solid == 0)[..., None].astype(float)
    d = (out[..., :nt] - target) * fluid
    count = fluid.sum() * nt
    loss = (d * d).sum() / count
    g = np.zeros_like(out)
    g[..., :nt] = 2.0 * d / count
    return loss, g


def backward(params, layers, skips, act, pre, g_out):
    grads = {}
    g = {"out": g_out}
    for name, cin, cout, stride, relu, src in reversed(layers):
        gz = g.pop(name)
        if relu:
            gz = gz * (pre[name] > 0)
        w, b = params[name]
        dx, dw, db = conv3x3_bwd(act[src], w, gz, stride)
        grads[name] = (dw, db)
        g[src] = g.get(src, 0.0) + dx
        if src in skips:
            a, bb = skips[src]
            gu = g.pop(src)
            g[a] = g.get(a, 0.0) + upsample_bwd(gu)
            g[bb] = g.get(bb, 0.0) + gu
    return grads
