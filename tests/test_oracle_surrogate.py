"""Pin the FP64 layer oracle of the surrogate by central finite differences."""

import numpy as np

from oracle import surrogate_np as S


def _num_grad(f, x, eps=1e-6):
    g = np.zeros_like(x)
    it = np.nditer(x, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        old = x[i]
        x[i] = old + eps
        fp = f()
        x[i] = old - eps
        fm = f()
        x[i] = old
        g[i] = (fp - fm) / (2 * eps)
    return g


def test_conv_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    for k, s, p, slope in ((3, 2, 1, 0.1), (3, 1, 1, 0.1), (2, 1, 1, 1.0)):
        x = rng.standard_normal((3, 2, 6, 5))
        w = rng.standard_normal((4, 3 * k * k)) * 0.3
        b = rng.standard_normal(4) * 0.1
        y, cache = S.conv_fwd(x, w, b, k, s, p, slope)
        r = rng.standard_normal(y.shape)
        loss = lambda: float((S.conv_fwd(x, w, b, k, s, p, slope)[0] * r).sum())
        dx, dw, db = S.conv_bwd(r, w, cache, k, s, p, slope)
        np.testing.assert_allclose(dw, _num_grad(loss, w), rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(db, _num_grad(loss, b), rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(dx, _num_grad(loss, x), rtol=1e-5, atol=1e-7)


def test_upsample_and_loss_gradients():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 3, 4))
    r = rng.standard_normal((2, 3, 6, 8))
    np.testing.assert_allclose(S.up2_bwd(r), _num_grad(lambda: float((S.up2_fwd(x) * r).sum()), x),
                               rtol=1e-6)
    y, t = rng.standard_normal((2, 3, 4, 4)), rng.standard_normal((2, 3, 4, 4))
    per, g = S.rel_sq_loss(y, t)
    np.testing.assert_allclose(g, _num_grad(lambda: float(S.rel_sq_loss(y, t)[0].mean()), y),
                               rtol=1e-5, atol=1e-8)
    assert per.shape == (3,)


def test_plane_bias_gradient():
    rng = np.random.default_rng(2)
    x, T = rng.standard_normal((2, 3, 4, 5)), rng.standard_normal((2, 4, 5))
    r = rng.standard_normal((2, 3, 4, 5))
    loss = lambda: float((S.plane_bias_fwd(x, T) * r).sum())
    np.testing.assert_allclose(S.plane_bias_bwd(r), _num_grad(loss, T), rtol=1e-6)
    np.testing.assert_allclose(S.plane_bias_fwd(x, T)[:, 1], x[:, 1] + T)


def test_energy_loss_gradient():
    rng = np.random.default_rng(3)
    y, t = rng.standard_normal((2, 4, 3, 3)), rng.standard_normal((2, 4, 3, 3))
    per, g = S.energy_sq_loss(y, t)
    np.testing.assert_allclose(g, _num_grad(lambda: float(S.energy_sq_loss(y, t)[0].mean()), y),
                               rtol=1e-5, atol=1e-8)
    # equal target energies: identical to the per-sample relative loss
    t1 = t / np.sqrt((t ** 2).sum((0, 2, 3)))[None, :, None, None]
    np.testing.assert_allclose(S.energy_sq_loss(y, t1)[0], S.rel_sq_loss(y, t1)[0])


def test_adam_first_step_closed_form():
    # the bias-corrected first step moves every weight by lr g / (|g| + eps) plus the decay
    rng = np.random.default_rng(4)
    w, g = rng.standard_normal(50), rng.standard_normal(50)
    w1, m, v = S.adam_step(w, g, np.zeros(50), np.zeros(50), 1, 0.01, eps=1e-8, wd=0.1)
    np.testing.assert_allclose(w1, w - 0.01 * (g / (np.abs(g) + 1e-8) + 0.1 * w), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_conv_forward_equals_direct_sum_and_library_correlation():
    """y[o, b, i, j] = leaky(b[o] + sum_{c, u, v} w[o, (c, u, v)] xpad[c, b, s i + u, s j + v]):
    a direct loop, and for stride 1 scipy's 2-D correlation per channel pair.  A flipped
    kernel, a transposed tap order or a wrong padding offset fails one of them."""
    from scipy.signal import correlate2d
    rng = np.random.default_rng(5)
    for k, s, p, slope in ((3, 2, 1, 0.1), (3, 1, 1, 0.2), (2, 1, 1, 1.0), (4, 2, 0, 0.05)):
        C, B, H, W, co = 3, 2, 7, 6, 4
        x = rng.standard_normal((C, B, H, W))
        w = rng.standard_normal((co, C * k * k))
        b = rng.standard_normal(co)
        y, _ = S.conv_fwd(x, w, b, k, s, p, slope)
        Ho, Wo = (H + 2 * p - k) // s + 1, (W + 2 * p - k) // s + 1
        assert y.shape == (co, B, Ho, Wo)
        xp = np.zeros((C, B, H + 2 * p, W + 2 * p))
        xp[:, :, p:p + H, p:p + W] = x
        wk = w.reshape(co, C, k, k)
        want = np.empty_like(y)
        for o, n, i, j in np.ndindex(co, B, Ho, Wo):
            z = b[o]
            for c, u, v in np.ndindex(C, k, k):
                z += wk[o, c, u, v] * xp[c, n, s * i + u, s * j + v]
            want[o, n, i, j] = z if z > 0 else slope * z
        np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-12)
        if s == 1:
            for o, n in np.ndindex(co, B):
                z = b[o] + sum(correlate2d(xp[c, n], wk[o, c], mode="valid") for c in range(C))
                np.testing.assert_allclose(y[o, n], np.where(z > 0, z, slope * z), rtol=1e-12,
                                           atol=1e-12)
