"""FP64 NumPy restatement of the surrogate network's layers -- TEST INFRASTRUCTURE ONLY.

Forward and backward of a k x k convolution (stride s, padding p) with bias
and leaky ReLU, nearest 2x upsampling, and the per-sample relative squared
error loss, on channel-major (C, B, H, W) arrays.  Used to check the GPU
kernels' gradients.  Pinned in tests/test_oracle_surrogate.py: the forward pass by a direct
sum and by scipy's 2-D correlation, the backward passes by finite differences.
"""

import numpy as np


def conv_fwd(x, w, b, k, s, p, slope):
    C, B, H, W = x.shape
    co = w.shape[0]
    Ho, Wo = (H + 2 * p - k) // s + 1, (W + 2 * p - k) // s + 1
    xp = np.pad(x, ((0, 0), (0, 0), (p, p), (p, p)))
    col = np.empty((C, k, k, B, Ho, Wo))
    for ky in range(k):
        for kx in range(k):
            col[:, ky, kx] = xp[:, :, ky:ky + s * Ho:s, kx:kx + s * Wo:s]
    z = w @ col.reshape(C * k * k, -1) + b[:, None]
    y = np.where(z > 0, z, slope * z).reshape(co, B, Ho, Wo)
    return y, (x.shape, col, z)


def conv_bwd(dy, w, cache, k, s, p, slope):
    (C, B, H, W), col, z = cache
    co = w.shape[0]
    g = dy.reshape(co, -1) * np.where(z > 0, 1.0, slope)
    dw = g @ col.reshape(C * k * k, -1).T
    db = g.sum(1)
    dcol = (w.T @ g).reshape(C, k, k, B, *dy.shape[2:])
    Ho, Wo = dy.shape[2:]
    dxp = np.zeros((C, B, H + 2 * p, W + 2 * p))
    for ky in range(k):
        for kx in range(k):
            dxp[:, :, ky:ky + s * Ho:s, kx:kx + s * Wo:s] += dcol[:, ky, kx]
    return dxp[:, :, p:p + H, p:p + W], dw, db


def up2_fwd(x):
    return x.repeat(2, axis=2).repeat(2, axis=3)


def up2_bwd(dy):
    C, B, H, W = dy.shape
    return dy.reshape(C, B, H // 2, 2, W // 2, 2).sum((3, 5))


def rel_sq_loss(y, t):
    """Per-sample |y - t|^2/|t|^2 and the gradient of their mean."""
    B = y.shape[1]
    d = y - t
    num = (d ** 2).sum((0, 2, 3))
    den = (t ** 2).sum((0, 2, 3))
    return num / den, 2 * d / den[None, :, None, None] / B


def energy_sq_loss(y, t):
    """Per-sample |y - t|^2 / E, E = mean over samples of |t|^2, and the gradient of the mean."""
    B = y.shape[1]
    d = y - t
    E = (t ** 2).sum((0, 2, 3)).mean()
    return (d ** 2).sum((0, 2, 3)) / E, 2 * d / E / B


def plane_bias_fwd(x, T):
    """Learned output planes: y[c, b] = x[c, b] + T[c] (channel-major (C, B, H, W))."""
    return x + T[:, None]


def plane_bias_bwd(g):
    """Gradient of the planes: sum over samples."""
    return g.sum(1)


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW step (bias-corrected moments, decoupled decay); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh, vh = m / (1 - b1 ** t), v / (1 - b2 ** t)
    return w - lr * (mh / (np.sqrt(vh) + eps) + wd * w), m, v
