"""CPU oracle of the learned corrector network (test infrastructure only).

Network: conv3x3(8->32) + ReLU, conv3x3(32->32) + ReLU, conv3x3(32->3), zero
padding, channels-last images [n][y][x][c]; flat parameter order
W1[32][8][3][3], b1, W2[32][32][3][3], b2, W3[3][32][3][3], b3.

Two arithmetic modes:
  exact=True   float64 throughout: the plain mathematics, pinned in the tests
               against finite differences and a loop-level convolution;
  exact=False  rounds to bf16 exactly where the HIP kernels do (input features,
               weights, hidden activations, the output gradient and the
               masked hidden gradients), accumulating in float64.
"""
from __future__ import annotations

import numpy as np

C_IN, C_H, C_OUT = 8, 32, 3
SHAPES = [("W1", (C_H, C_IN, 3, 3)), ("b1", (C_H,)), ("W2", (C_H, C_H, 3, 3)), ("b2", (C_H,)),
          ("W3", (C_OUT, C_H, 3, 3)), ("b3", (C_OUT,))]


def n_params():
    return sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, k = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[k:k + n], dtype=np.float64).reshape(shp)
        k += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None):
    """y[n,i,j,o] = sum_{c,ky,kx} w[o,c,ky,kx] x[n,i+ky-1,j+kx-1,c] (+ b[o])."""
    n, H, W, C = x.shape
    xp = np.zeros((n, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1, :] = x
    y = np.zeros((n, H, W, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("nijc,oc->nijo", xp[:, ky:ky + H, kx:kx + W, :], w[:, :, ky, kx])
    if b is not None:
        y += b
    return y


def conv3x3_wgrad(x, dy):
    """dW[o,c,ky,kx] = sum_{n,i,j} dy[n,i,j,o] x[n,i+ky-1,j+kx-1,c]."""
    n, H, W, C = x.shape
    xp = np.zeros((n, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1, :] = x
    g = np.zeros((dy.shape[-1], C, 3, 3))
    for ky in range(3):
        for kx in range(3):
            g[:, :, ky, kx] = np.einsum("nijo,nijc->oc", dy, xp[:, ky:ky + H, kx:kx + W, :])
    return g


def conv3x3_dgrad(dy, w):
    """dx = transpose of conv3x3 applied to dy (mirror taps, swap channels)."""
    wt = np.transpose(w[:, :, ::-1, ::-1], (1, 0, 2, 3))
    return conv3x3(dy, wt)


def features(raw, shift, scale, exact):
    x = (np.asarray(raw, dtype=np.float32) - np.float32(1) * np.asarray(shift, np.float32)) * np.asarray(scale, np.float32)
    return np.asarray(x, dtype=np.float64) if exact else bf16(x)


def forward(flat, raw, shift, scale, exact=False):
    p = unpack(flat)
    r = (lambda a: a) if exact else bf16
    x0 = features(raw, shift, scale, exact)
    a1 = r(np.maximum(conv3x3(x0, r(p["W1"]), p["b1"]), 0))
    a2 = r(np.maximum(conv3x3(a1, r(p["W2"]), p["b2"]), 0))
    y = conv3x3(a2, r(p["W3"]), p["b3"])
    return y, (x0, a1, a2)


def band_mask(H, W, band):
    """1 on the points at least `band` from every wall of an H x W image, else
    0 (the corrected region)."""
    dx = np.minimum(np.arange(W), W - 1 - np.arange(W))
    dy = np.minimum(np.arange(H), H - 1 - np.arange(H))
    return (np.minimum(dy[:, None], dx[None, :]) >= band).astype(np.float64)


def loss_and_grads(flat, raw, target, shift, scale, out_scale, exact=False, band=0):
    """MSE of (net - target/out_scale) over the points of the corrected
    region (band_mask) and all channels, divided by the count of every point
    and channel, and the gradient of that loss with respect to every
    parameter (flat order)."""
    p = unpack(flat)
    r = (lambda a: a) if exact else bf16
    y, (x0, a1, a2) = forward(flat, raw, shift, scale, exact)
    t = np.asarray(target, dtype=np.float64) / np.asarray(out_scale, dtype=np.float64)
    if not exact:
        t = np.asarray(np.asarray(target, np.float32) / np.asarray(out_scale, np.float32), np.float64)
    N = y.size
    d = (y - t) * band_mask(y.shape[1], y.shape[2], band)[None, :, :, None]
    loss = float(np.sum(d * d) / N)
    dy = r(2.0 * d / N)
    g = {"W3": conv3x3_wgrad(a2, dy), "b3": dy.sum(axis=(0, 1, 2))}
    da2 = conv3x3_dgrad(dy, r(p["W3"]))
    dp2 = r(da2 * (a2 > 0))
    g["W2"] = conv3x3_wgrad(a1, dp2)
    g["b2"] = dp2.sum(axis=(0, 1, 2))
    da1 = conv3x3_dgrad(dp2, r(p["W2"]))
    dp1 = r(da1 * (a1 > 0))
    g["W1"] = conv3x3_wgrad(x0, dp1)
    g["b1"] = dp1.sum(axis=(0, 1, 2))
    return loss, pack(g)
