"""Pin the input-generation restatement (oracle/inputs_np.py) by closed forms and brute force."""

import numpy as np
import pytest

from oracle import inputs_np as I
from oracle import OracleConfig


def _plane_wave(dims, j, amp, imag=False):
    """Inverse DFT of one spectral sample placed at voxel j, read back through the two
    centring shifts: the spectrum index of j is (j - n//2) mod n and output voxel i reads
    transform index (i + n//2) mod n."""
    phase = np.zeros(dims)
    for d, n in enumerate(dims):
        jd = (j[d] - n // 2) % n
        idx = (np.arange(n) + n // 2) % n
        shape = [1, 1, 1]
        shape[d] = n
        phase = phase + (2 * np.pi * jd * idx / n).reshape(shape)
    return -amp * np.sin(phase) if imag else amp * np.cos(phase)


@pytest.mark.parametrize("dims,j", [((7, 5, 3), (4, 2, 1)), ((7, 5, 3), (0, 4, 2)),
                                    ((6, 4, 5), (1, 3, 0)), ((8, 9, 2), (5, 0, 1))])
@pytest.mark.parametrize("imag", [False, True])
def test_single_frequency_is_closed_form_plane_wave(dims, j, imag):
    a, b, c, w, p = 1.3, 0.7, 2.1, 0.6, 5
    re, im = np.zeros(dims), np.zeros(dims)
    (im if imag else re)[j] = 1.0
    got = I.masked_gp(re, im, [[a, b, c]], p, l_weights=[w])
    N = np.prod(dims)
    env = np.sqrt(a * b * c) * w
    for d, (l, n) in enumerate(zip((a, b, c), dims)):
        g = (-np.pi / 2 + np.pi * j[d] / (n - 1)) ** 2
        env *= np.exp(-g * l * l)
    want = _plane_wave(dims, j, 2 * p * np.sqrt(N) * env / N, imag)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_envelopes_of_scales_add_in_power():
    dims, j = (9, 7, 5), (6, 2, 3)
    re, im = np.zeros(dims), np.zeros(dims)
    re[j] = 1.0
    s1, s2 = [0.9, 1.4, 0.3], [2.2, 0.5, 1.1]
    one = np.abs(I.masked_gp(re, im, [s1], 1, l_weights=[0.8])).max()
    two = np.abs(I.masked_gp(re, im, [s2], 1, l_weights=[0.3])).max()
    both = np.abs(I.masked_gp(re, im, [s1, s2], 1, l_weights=[0.8, 0.3])).max()
    assert both == pytest.approx(np.hypot(one, two), rel=1e-12)


def test_field_variance_is_four_p_squared_mean_envelope_power():
    dims, ls, p = (48, 40, 16), [[1.5, 1.5, 0.6], [4.0, 4.0, 1.0]], 3
    env2 = (I.envelope(dims, ls, [0.6, 0.3]) ** 2).mean()
    var = []
    for seed in range(4):
        rng = np.random.default_rng(seed)
        f = I.masked_gp(rng.standard_normal(dims), rng.standard_normal(dims), ls, p,
                        l_weights=[0.6, 0.3])
        var.append(np.mean(f ** 2))
    assert np.mean(var) == pytest.approx(4 * p * p * env2, rel=0.06)


def _mask_cfg():
    # r = 1, pc = 5, collection radius 0.7346 * (8 - depth): slab 0 radius 4.04, slab 1 0.37
    # (the centre pixel only), deeper slabs skipped
    return OracleConfig(volume_um=(8.0, 8.0, 6.0), depth_um=5.0, tail_um=2.0, psf_um=(4, 4, 4))


def test_uniform_vessels_give_closed_form_absorption():
    cfg = _mask_cfg()
    m = I.collection_mask(np.ones((40, 40, 16)), cfg)
    assert m.shape == (8, 8)
    np.testing.assert_allclose(m, 10 ** (-10.0 * cfg.hemoabs), rtol=1e-12)


def test_single_vessel_voxel_gives_disk_footprint():
    cfg = _mask_cfg()
    v = np.zeros((40, 40, 16))
    v[19, 21, 2] = 1.0
    total = -np.log10(I.collection_mask(v, cfg)) * cfg.vres / cfg.hemoabs
    radius = cfg.vres * np.tan(np.arcsin(cfg.objNA / cfg.n)) * (
        cfg.depth_um + cfg.volume_um[2] / 2 - 0.5 * cfg.prop_um)
    ii, jj = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    count = sum(1 for dx in range(-6, 7) for dy in range(-6, 7) if dx * dx + dy * dy <= radius ** 2)
    foot = ((ii - 19) ** 2 + (jj - 21) ** 2 <= radius ** 2) / count
    np.testing.assert_allclose(total, foot[16:24, 16:24], atol=1e-12)


def test_disk_convolution_equals_direct_sum_on_odd_grid():
    # odd grid sides give half-integer disk coordinates and even disk sizes
    from scipy.signal import convolve2d
    cfg = OracleConfig(volume_um=(9.0, 7.0, 6.0), depth_um=9.0, tail_um=2.0, psf_um=(4, 4, 4))
    rng = np.random.default_rng(3)
    v = (rng.random((21, 19, 12)) < 0.1).astype(float)
    got = -np.log10(I.collection_mask(v, cfg)) / cfg.hemoabs
    total = np.zeros((21, 19))
    for s in range(0, 12, 5):
        rad = np.tan(np.arcsin(cfg.objNA / cfg.n)) * (cfg.depth_um + 3 - (s / 5 + 0.5) * 5)
        if rad <= 0:
            continue
        x = np.arange(-10.5, 10.5)[:, None]
        y = np.arange(-9.5, 9.5)[None, :]
        d = (x * x + y * y <= rad * rad).astype(float)
        rows, cols = np.flatnonzero(d.any(1)), np.flatnonzero(d.any(0))
        d = d[rows[0]:rows[-1] + 1, cols[0]:cols[-1] + 1]
        total += convolve2d(v[..., s:s + 5].sum(-1), d / d.sum(), mode="same")
    np.testing.assert_allclose(got, total[6:15, 6:13], atol=1e-10)


def test_partial_slabs_and_occupancy_mix():
    # za = 6 and nc = 2 with pc = 5: both end slabs are partial
    cfg = _mask_cfg()
    shapes = I.field_shapes(cfg)
    rng = np.random.default_rng(1)
    noises = [(rng.standard_normal(s), rng.standard_normal(s)) for s in shapes]
    base = I.prepare_fields(cfg, noises)
    full = [I.masked_gp(re, im, np.repeat(np.array(cfg.scatter_scales_um)[:, None], 3, 1)
                        * [1, 1, 1.0 / (1 if i == 1 else 5)], 1 if i == 1 else 5,
                        l_weights=cfg.scatter_weights) for i, (re, im) in enumerate(noises)]
    np.testing.assert_allclose(base[0][..., 0], full[0][..., 0] * 0.2, rtol=1e-12)
    np.testing.assert_allclose(base[0][..., 1:], full[0][..., 1:], rtol=1e-12)
    np.testing.assert_allclose(base[2][..., -1], full[2][..., -1] * 0.4, rtol=1e-12)
    np.testing.assert_allclose(base[1], full[1], rtol=1e-12)
    # occupied voxels: occ + raw (1 - occ / slab); a whole occupied slab is pure occupancy
    X, Y = shapes[0][:2]
    vessels = np.ones((X, Y, 6 + 4 + 2))
    mixed = I.prepare_fields(cfg, noises, vessels=vessels)
    np.testing.assert_allclose(mixed[0][..., 0], 1 + base[0][..., 0] * 0.8, rtol=1e-12)
    np.testing.assert_allclose(mixed[0][..., 1], 5.0, atol=1e-12)
    np.testing.assert_allclose(mixed[1], 1.0, atol=1e-12)
    np.testing.assert_allclose(mixed[2][..., 0], 2 + base[2][..., 0] * 0.6, rtol=1e-12)
