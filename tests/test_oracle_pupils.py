"""Pins of the pupil oracle (oracle/pupils_np.py; maomi/pupils.py of the reference).

The reference's tests hold no vectors for these fields, so each check is a closed form, a
special case that reduces to a golden-pinned function, or a symmetry, chosen so that a dropped
factor, a wrong sign or a swapped axis fails one of them.
"""

import numpy as np
import pytest

import oracle as O
from oracle import pupils_np as PO

DEMO = dict(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)


@pytest.fixture(params=["default", "demo"])
def cfg(request):
    return O.OracleConfig() if request.param == "default" else O.OracleConfig(**DEMO)


def test_gaussian_kind_is_generate_ba(cfg):
    """make_pupil's gaussian kind is generate_ba's pupil (golden-pinned) up to the float32
    rounding of the radii and the focus coefficient; the focus phase reaches ~1e4 rad, so the
    phase bar is that rounding times the phase."""
    a = PO.make_pupil(cfg)
    b = O.generate_ba(cfg)
    assert a.shape == b.shape
    np.testing.assert_allclose(np.abs(a), np.abs(b), rtol=0, atol=2e-6)
    on = np.abs(b) > 1e-3
    dphi = np.angle(a[on] * np.conj(b[on]))
    g = PO.pupil_grid(cfg)
    theta = np.abs(g.k / (2 * g.fl) * (g.x ** 2 + g.y ** 2))[on]
    assert np.all(np.abs(dphi) <= 4e-7 * theta + 1e-4)


def test_cylindrical_kind(cfg):
    g = PO.pupil_grid(cfg)
    ref = PO.make_pupil(cfg)
    np.testing.assert_array_equal(PO.make_pupil(cfg, kind="cylindrical"), ref)
    sx, sy = 1.7, 0.45
    cyl = PO.make_pupil(cfg, kind="cylindrical", cylscale=(sx, sy))
    x, y = g.x.astype(float), g.y.astype(float)
    ellipse = (x / sx) ** 2 + (y / sy) ** 2 < float(g.objrad) ** 2
    np.testing.assert_array_equal(cyl, np.where(ellipse, ref, 0))
    assert 0 < ellipse.sum() < (np.abs(ref) > 0).sum()
    for bad in [(1.0,), (0.0, 1.0), (np.inf, 1.0), (1.0, -2.0)]:
        with pytest.raises(ValueError, match="cylindrical scale"):
            PO.make_pupil(cfg, kind="cylindrical", cylscale=bad)


@pytest.mark.parametrize("kind", ["gaussian", "bessel"])
def test_pair_at_zero_angle_and_separation_is_single_beam(cfg, kind):
    g = PO.pupil_grid(cfg)
    beam_na = float(cfg.NA)
    pair = PO.make_vtwins_pupils(cfg, beam_na=beam_na, theta_deg=0.0, separation_um=0.0,
                                 kind=kind, length_um=30.0)
    single = PO.make_pupil(cfg, kind=kind, length_um=15.0)  # the pair's width omits the 2
    if kind == "gaussian":
        np.testing.assert_array_equal(pair["left"], single)
    else:  # the single Bessel beam is not clipped to the objective
        inside = g.x * g.x + g.y * g.y < g.objrad * g.objrad
        np.testing.assert_array_equal(pair["left"], np.where(inside, single, 0))
    np.testing.assert_array_equal(pair["right"], pair["left"])


def test_pair_beams_mirror_in_x_and_tilt_in_opposite_directions(cfg):
    g = PO.pupil_grid(cfg)
    n0 = len(g.xs)
    assert np.array_equal(g.xs[1:][::-1], -g.xs[1:])  # x_{n-j} = -x_j for j >= 1
    tilted = PO.make_vtwins_pupils(cfg, theta_deg=24.0, separation_um=6.0)
    flat = PO.make_vtwins_pupils(cfg, theta_deg=24.0, separation_um=0.0)
    left, right = np.abs(flat["left"]), np.abs(flat["right"])
    np.testing.assert_allclose(left[:, 1:], right[:, 1:][:, ::-1], rtol=1e-6, atol=1e-12)
    # the beams sit on opposite sides of the axis: shift = tan(theta/2) fl
    cx = (left * g.x).sum() / left.sum()
    assert np.isclose(cx, np.tan(np.deg2rad(12.0)) * float(g.fl), rtol=2e-2)
    # the tilt is a linear phase of slope k * lambda * separation / (4 aperture), mean zero
    ph = PO.tilt_phase(cfg, g, 6.0, g.objrad)
    slope = np.polyfit(g.xs.astype(float), ph[0].astype(float), 1)[0]
    assert np.isclose(slope, cfg.wavelength_um * 1e-6 * 6.0 / 4 / float(g.objrad), rtol=1e-5)
    assert abs(float(ph.astype(float).mean())) < 1e-6 * float(np.abs(ph).max())
    for label, sign in (("left", 1), ("right", -1)):
        a, b = tilted[label], flat[label]
        on = np.abs(b) > 1e-3 * np.abs(b).max()
        np.testing.assert_allclose(np.abs(a[on]), np.abs(b[on]), rtol=1e-5)
        d = np.angle(a[on] * np.conj(b[on]) * np.exp(-1j * sign * float(g.k) * ph[on]))
        assert np.abs(d).max() < 1e-3
    assert n0 == len(g.xs)


def test_ring_profiles():
    w = np.float32(2.0e-4)
    dr = np.linspace(-8e-4, 8e-4, 4001).astype(np.float32)
    gauss = PO.ring_amplitude(dr, w, "gaussian")
    assert gauss[2000] == 1.0 and dr[2000] == 0.0
    half = PO.ring_amplitude(np.float32([w / 2, -w / 2]), w, "gaussian")
    np.testing.assert_allclose(half, 0.5, rtol=1e-6)  # full width at half maximum = w
    uni = PO.ring_amplitude(dr, w, "uniform")
    np.testing.assert_array_equal(uni, np.abs(dr.astype(float)) < float(w) / 2)
    pos = PO.ring_amplitude(dr, w, "posax")
    neg = PO.ring_amplitude(dr[::-1], w, "negax")
    np.testing.assert_array_equal(pos, neg)  # negax(dr) = posax(-dr)
    wp = float(w) / 1.133
    assert abs(float(dr[np.argmax(pos)])) <= 1e-6  # z e^{-(z/w)^2} peaks at z = w/sqrt2, dr = 0
    np.testing.assert_allclose(pos.max(), wp / np.sqrt(2) * np.exp(-0.5), rtol=1e-5)
    assert np.all(pos[dr.astype(float) <= -wp / np.sqrt(2) - 1e-9] == 0)
    with pytest.raises(ValueError, match="Unknown Bessel profile"):
        PO.ring_amplitude(dr, w, "cosine")


def test_bessel_ring_sits_on_the_na_radius(cfg):
    g = PO.pupil_grid(cfg)
    f = PO.make_pupil(cfg, kind="bessel", profile="uniform", length_um=20.0)
    r = np.sqrt(g.x.astype(float) ** 2 + g.y.astype(float) ** 2)
    width = 1.4e-10 * np.tan(np.arcsin(cfg.NA / cfg.n)) / (float(g.fl) * 1e-6 * 2 * 20.0)
    on = np.abs(f) > 0
    assert np.all(np.abs(r[on] - float(g.rad)) < width / 2 * (1 + 1e-5))
    d1 = float(g.xs[1] - g.xs[0])
    area = 2 * np.pi * float(g.rad) * width / d1 ** 2
    assert abs(on.sum() - area) < 0.1 * area + 8


def test_rejections(cfg):
    with pytest.raises(ValueError, match="single-field kinds"):
        PO.make_pupil(cfg, kind="airy")
    with pytest.raises(ValueError, match="Bessel length"):
        PO.make_pupil(cfg, kind="bessel", length_um=0.0)
    for kw in [dict(beam_na=0.0), dict(beam_na=2.0), dict(theta_deg=180.0),
               dict(length_um=-1.0), dict(separation_um=np.nan)]:
        with pytest.raises(ValueError, match="Invalid Vtwins"):
            PO.make_vtwins_pupils(cfg, **kw)
    with pytest.raises(ValueError, match="gaussian or bessel"):
        PO.make_vtwins_pupils(cfg, kind="cylindrical")
