"""2-D PSF contracts (maomi/psf.py:13-203): validation, NPZ round trips, provider sampling,
and the oracle of resample_psf / pupil_to_intensity pinned by closed forms (CPU only)."""

import json

import numpy as np
import pytest

from maomi_amd.psf import PSFField, SampledPSF, pupil_to_intensity, resample_psf
from oracle import psf_np as OP


def _kernel(shape=(5, 7), seed=0):
    a = np.random.default_rng(seed).random(shape) + 0.01
    return a / a.sum()


# ------------------------------------------------------------------------ SampledPSF
def test_sampled_psf_validation_texts():
    k = _kernel()
    p = SampledPSF(k, (0.5, 0.25), 920, {"a": 1})
    assert p.pixel_size_um == (0.5, 0.25) and not p.intensity.flags.writeable
    assert p.provenance == {"a": 1}
    cases = [
        (dict(intensity=np.ones(5) / 5), "finite, nonnegative 2-D"),
        (dict(intensity=np.ones((1, 5)) / 5), "finite, nonnegative 2-D"),
        (dict(intensity=-k), "finite, nonnegative 2-D"),
        (dict(intensity=k * 2), "unit sum"),
        (dict(pixel_size_um=(0.5,)), "two positive specimen-plane pixel pitches in micrometres"),
        (dict(pixel_size_um=(0.5, 0)), "two positive"),
        (dict(wavelength_nm=0), "wavelength_nm must be positive and finite"),
        (dict(wavelength_nm=np.inf), "wavelength_nm must be positive and finite"),
    ]
    for kw, msg in cases:
        args = dict(intensity=k, pixel_size_um=(0.5, 0.5), wavelength_nm=920.0)
        args.update(kw)
        with pytest.raises(ValueError, match=msg):
            SampledPSF(**args)
    with pytest.raises(ValueError):
        SampledPSF(k, (1, 1), 920, {"bad": float("nan")})
    with pytest.raises(ValueError, match="Cannot normalize"):
        SampledPSF.normalized(np.zeros((3, 3)), (1, 1), 920)
    n = SampledPSF.normalized(k * 7, (1, 1), 920)
    np.testing.assert_allclose(n.intensity, k, rtol=1e-15)


def test_sampled_psf_npz_schema_round_trip(tmp_path):
    p = SampledPSF(_kernel(), (0.5, 0.25), 920.0, {"src": "x", "n": [1, 2]})
    path = tmp_path / "sub" / "k.npz"
    p.save(path)
    with np.load(path, allow_pickle=False) as d:  # the reference's keys and values
        assert sorted(d.files) == ["axes", "intensity", "normalization", "pixel_size_um",
                                   "provenance", "wavelength_nm"]
        assert str(d["axes"]) == "XY" and str(d["normalization"]) == "unit_sum"
        assert json.loads(str(d["provenance"])) == {"src": "x", "n": [1, 2]}
    q = SampledPSF.load(path)
    np.testing.assert_array_equal(q.intensity, p.intensity)
    assert q.pixel_size_um == p.pixel_size_um and q.wavelength_nm == 920.0
    np.savez(tmp_path / "bad.npz", intensity=p.intensity, pixel_size_um=(1, 1), wavelength_nm=1,
             axes="YX", normalization="unit_sum", provenance="{}")
    with pytest.raises(ValueError, match="Unsupported PSF axes or normalization"):
        SampledPSF.load(tmp_path / "bad.npz")


# ------------------------------------------------------------------------ PSFField
def _provider(shape=(5, 5), pitch=(0.5, 0.5), wavelength=920.0):
    calls = []

    def provider(x, y):
        calls.append((x, y))
        g = np.exp(-((np.arange(shape[0])[:, None] - shape[0] // 2 - 0.1 * x) ** 2
                     + (np.arange(shape[1])[None, :] - shape[1] // 2 - 0.1 * y) ** 2))
        return SampledPSF.normalized(g, pitch, wavelength, {"x": x, "y": y})
    return provider, calls


def test_psf_field_from_provider_samples_each_node_once():
    prov, calls = _provider()
    x, y = [0.0, 1.0, 3.0], [2.0, 5.0]
    f = PSFField.from_provider(prov, x, y)
    assert calls == [(a, b) for a in x for b in y]
    assert f.intensity.shape == (3, 2, 5, 5)
    for i, a in enumerate(x):
        for j, b in enumerate(y):
            np.testing.assert_array_equal(f.intensity[i, j], prov(a, b).intensity)
    assert f.provenance["interpolation"] == "bilinear at source; nearest edge extension"
    assert f.provenance["samples"][1] == {"x_um": 0.0, "y_um": 5.0, "provider": {"x": 0.0, "y": 5.0}}
    assert not f.intensity.flags.writeable and not f.x_um.flags.writeable


def test_psf_field_provider_errors_and_prefetch():
    with pytest.raises(TypeError, match="must return SampledPSF"):
        PSFField.from_provider(lambda x, y: np.ones((3, 3)) / 9, [0.0], [0.0])
    a, _ = _provider()
    b, _ = _provider(shape=(7, 5))
    mixed = lambda x, y: (a if x < 1 else b)(x, y)
    with pytest.raises(ValueError, match="share shape, pitch and wavelength; resample explicitly"):
        PSFField.from_provider(mixed, [0.0, 2.0], [0.0])
    c, _ = _provider(wavelength=1000.0)
    with pytest.raises(ValueError, match="share shape, pitch and wavelength"):
        PSFField.from_provider(lambda x, y: (a if x < 1 else c)(x, y), [0.0, 2.0], [0.0])
    for bad in ([0.0, 0.0], [[0.0, 1.0]], [], [0.0, np.nan]):
        with pytest.raises(ValueError, match="finite and strictly increasing"):
            PSFField.from_provider(a, bad, [0.0])

    class Batched:
        def __init__(self):
            self.grids, self.inner = [], _provider()[0]

        def prefetch(self, x, y):
            self.grids.append((list(x), list(y)))

        def __call__(self, x, y):
            return self.inner(x, y)
    bp = Batched()
    PSFField.from_provider(bp, [0.0, 1.0], [0.0, 2.0, 4.0])
    assert bp.grids == [([0.0, 1.0], [0.0, 2.0, 4.0])]


def test_psf_field_validation_and_round_trip(tmp_path):
    prov, _ = _provider()
    f = PSFField.from_provider(prov, [0.0, 1.0], [0.0])
    with pytest.raises(ValueError, match="field_X, field_Y, kernel_X, kernel_Y"):
        PSFField(f.intensity[0], f.x_um, f.y_um, (0.5, 0.5), 920.0)
    with pytest.raises(ValueError, match="unit sum"):
        PSFField(f.intensity * 2, f.x_um, f.y_um, (0.5, 0.5), 920.0)
    with pytest.raises(ValueError, match="strictly increasing"):
        PSFField(f.intensity, [1.0, 0.0], f.y_um, (0.5, 0.5), 920.0)
    path = tmp_path / "f.npz"
    f.save(path)
    with np.load(path, allow_pickle=False) as d:
        assert str(d["axes"]) == "field_X,field_Y,kernel_X,kernel_Y"
        assert str(d["interpolation"]) == "source_bilinear_edge"
        assert str(d["normalization"]) == "unit_sum"
    g = PSFField.load(path)
    np.testing.assert_array_equal(g.intensity, f.intensity)
    np.testing.assert_array_equal(g.x_um, f.x_um)
    assert g.provenance == f.provenance


def test_gpu_functions_validate_before_touching_the_device():
    k = SampledPSF(_kernel(), (0.5, 0.5), 920.0)
    with pytest.raises(ValueError, match="two positive specimen-plane pixel pitches"):
        resample_psf(k, (0.5, -1), (5, 5))
    with pytest.raises(ValueError, match="two integer dimensions >= 2"):
        resample_psf(k, (0.5, 0.5), (5, 1))
    with pytest.raises(ValueError, match="OPD must be a 2-D grid"):
        pupil_to_intensity(np.zeros(5))
    with pytest.raises(ValueError, match="no smaller than the pupil grid"):
        pupil_to_intensity(np.zeros((9, 9)), fft_size=8)
    with pytest.raises(ValueError, match="must match the OPD"):
        pupil_to_intensity(np.zeros((4, 4)), 8, amplitude=np.ones((3, 4)))
    with pytest.raises(ValueError, match="No transmitting pupil samples"):
        pupil_to_intensity(np.full((4, 4), np.nan), 8)


# ------------------------------------------------------------------------ oracle pins
def test_oracle_resample_identity_ramp_and_mass():
    k = _kernel((7, 6), seed=3)
    same, mass = OP.resample_linear(k, (0.5, 0.25), (0.5, 0.25), (7, 6))
    np.testing.assert_allclose(same, k, rtol=1e-14)
    assert abs(mass - 1) < 1e-14
    # pitches that are not binary fractions round the edge coordinates (i - n//2) step / old:
    # (-3 * 0.4) / 0.4 + 3 = -4.4e-16 lies outside the support, so the first row and column are 0
    edge, _ = OP.resample_linear(k, (0.4, 0.4), (0.4, 0.4), (7, 6))
    assert (edge[:, 0] == 0).all() and (edge[0] == 0).all() and (edge[1:, 1:] > 0).all()
    # a linear ramp is reproduced wherever the sample lies inside the support, zero elsewhere
    from scipy.ndimage import map_coordinates
    ramp = 1.0 + np.add.outer(np.arange(9) * 0.3, np.arange(8) * 0.7)
    cx, cy = OP.resample_coordinates((15, 14), (0.35, 0.3), (0.5, 0.5), ramp.shape)
    raw = map_coordinates(ramp, np.array(np.meshgrid(cx, cy, indexing="ij")), order=1,
                          mode="constant", prefilter=False)
    inside = ((cx >= 0) & (cx <= 8))[:, None] & ((cy >= 0) & (cy <= 7))[None, :]
    want = 1.0 + 0.3 * cx[:, None] + 0.7 * cy[None, :] + 0 * raw
    np.testing.assert_allclose(raw[inside], want[inside], rtol=1e-13)
    assert inside.any() and (~inside).any() and (raw[~inside] == 0).all()
    # halving the pitch on a grid twice as large keeps the mass (4 samples per source pixel
    # at a quarter of the area each)
    big = np.zeros((9, 9))
    big[2:7, 2:7] = _kernel((5, 5), seed=4)
    _, m = OP.resample_linear(big, (0.25, 0.25), (0.5, 0.5), (18, 18))
    assert abs(m - 1) < 0.05


def test_oracle_pupil_closed_forms():
    # zero OPD on an s x s square: separable Dirichlet kernel sin^2(pi s k / n)/sin^2(pi k / n)
    for s, n in ((4, 32), (5, 33), (7, 64)):
        I = OP.pupil_to_intensity(np.zeros((s, s)), n)
        assert abs(I.sum() - 1) < 1e-13
        assert np.unravel_index(np.argmax(I), I.shape) == (n // 2, n // 2)
        k = np.arange(n) - n // 2
        with np.errstate(divide="ignore", invalid="ignore"):
            d = np.where(k == 0, s * s, np.sin(np.pi * s * k / n) ** 2 / np.sin(np.pi * k / n) ** 2)
        want = np.outer(d, d)
        np.testing.assert_allclose(I, want / want.sum(), rtol=1e-10, atol=1e-16)
    # a tilt of m / n waves per pupil sample moves the peak by m pixels along that axis
    n, s = 64, 12
    for m in (3, -5):
        opd = np.tile((np.arange(s) - s // 2)[:, None] * m / n, (1, s))
        I = OP.pupil_to_intensity(opd, n)
        assert np.unravel_index(np.argmax(I), I.shape) == (n // 2 + m, n // 2)
    # blocked (NaN) samples and zero amplitude transmit nothing: same as a smaller aperture
    opd = np.full((6, 6), np.nan)
    opd[1:5, 1:5] = 0.0
    amp = np.ones((6, 6))
    amp[0] = 0.0
    np.testing.assert_allclose(OP.pupil_to_intensity(opd, 32, amp),
                               OP.pupil_to_intensity(np.zeros((4, 4)), 32), atol=1e-15)
