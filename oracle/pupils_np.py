"""Alternative input pupils (maomi/pupils.py) in NumPy float32 -- TEST INFRASTRUCTURE ONLY.

The reference builds these fields with float32 NumPy array arithmetic; this module restates
that arithmetic step by step with the same float32 scalars, so that it is the checker for
maomi_pupil_field.  Pinned in tests/test_oracle_pupils.py (the reference's own tests hold no
vectors for this module):
- make_pupil's gaussian kind against generate_ba (golden-pinned): same field to float32 rounding
  of the radius;
- the cylindrical kind at unit scales against the gaussian kind, bit for bit, and its support
  against the closed-form ellipse;
- a pair beam at zero angle and zero separation against make_pupil's gaussian at the beam NA,
  bit for bit; the two beams' amplitudes as mirror images in x; the tilt as a linear phase of
  the stated slope;
- each ring profile against its closed form at dr = 0, its half maximum, its support and its
  mirror (posax against negax).
"""

from types import SimpleNamespace

import numpy as np

from .optics_np import gaussian_beam_size, vasc_size

F = np.float32


def pupil_grid(config):
    """pupils.py:37-51.  Returns the float32 grid and scalars: x, y (n1 x n0 meshgrid, x along
    columns), rad, objrad, k, fl, nr, and the 1-D coordinate vectors xs, ys."""
    fl = F(config.focal_length_mm / 1000)
    d2 = F(1e-6 / config.vres / config.ss)
    n = F(1e-6 * (vasc_size(config)[:2] - np.asarray(config.volume_um, float)[:2]) / d2)
    beam = gaussian_beam_size(config.objNA, config.n, float(fl * 1e6))
    d1 = F(max(beam) / 1e6 / min(n))
    nr = F(config.n)
    rad = F(np.tan(np.arcsin(F(config.NA) / nr)) * fl)
    objrad = F(np.tan(np.arcsin(F(config.objNA) / nr)) * fl)
    k = F(2 * nr * np.pi / F(config.wavelength_um * 1e-6))
    xs = np.arange(-n[0] / 2, n[0] / 2, dtype=F) * d1
    ys = np.arange(-n[1] / 2, n[1] / 2, dtype=F) * d1
    x, y = np.meshgrid(xs, ys)
    return SimpleNamespace(x=x, y=y, xs=xs, ys=ys, rad=rad, objrad=objrad, k=k, fl=fl, nr=nr)


def focus_phase(g):
    """exp(-i k rho^2 / (2 fl)) as the reference forms it (pupils.py:10, 81)."""
    rho2 = g.x * g.x + g.y * g.y
    return np.exp(np.complex64(-1j) * g.k / (2 * g.fl) * rho2)


def ring_amplitude(dr, width, profile):
    """pupils.py:112-122: the Bessel ring's radial profile at offset dr from the ring."""
    if profile == "gaussian":
        w = width / (2 * np.sqrt(F(np.log(2))))
        return np.exp(-((dr / w) ** 2))
    if profile == "uniform":
        return (dr < width / 2) & (dr > -width / 2)
    if profile in ("posax", "negax"):
        w = width / 1.133
        z = (dr if profile == "posax" else -dr) + w / np.sqrt(F(2))
        return z * np.exp(-((z / w) ** 2)) * (z > 0)
    raise ValueError("Unknown Bessel profile")


def make_pupil(config, *, kind="gaussian", length_um=60.0, profile="gaussian",
               cylscale=(1.0, 1.0)):
    """pupils.py:7-34."""
    g = pupil_grid(config)
    rho2 = g.x * g.x + g.y * g.y
    inside = rho2 < g.objrad * g.objrad
    if kind == "gaussian":
        amp = np.exp(-rho2 / (g.rad * g.rad)) * inside
    elif kind == "cylindrical":
        s = np.asarray(cylscale, float)
        if s.shape != (2,) or not np.isfinite(s).all() or np.any(s <= 0):
            raise ValueError("Positive two-axis cylindrical scale required")
        ell = (g.x / s[0]) ** 2 + (g.y / s[1]) ** 2
        support = (np.exp(-ell / (g.rad * g.rad)) * (ell < g.objrad * g.objrad)) > 0
        amp = np.exp(-rho2 / (g.rad * g.rad)) * inside * support
    elif kind == "bessel":
        if not np.isfinite(length_um) or length_um <= 0:
            raise ValueError("Positive Bessel length required")
        width = F(1.4e-10 * np.tan(np.arcsin(F(config.NA) / g.nr)) / (g.fl * 1e-6 * 2 * length_um))
        amp = ring_amplitude(np.sqrt(rho2) - g.rad, width, profile)
    else:
        raise ValueError("Supported single-field kinds: gaussian, bessel, cylindrical")
    field = np.asarray(amp * focus_phase(g), np.complex64)
    if not np.isfinite(field).all() or not np.any(field):
        raise ValueError("Pupil has no finite nonzero energy on this grid")
    return field


def tilt_phase(config, g, separation_um, aperture):
    """pupils.py:99-105: the pair's linear phase along x, mean removed (float32 mean over the
    whole grid)."""
    ph = (F(config.wavelength_um * 1e-6 * separation_um / 8) * F(2) * g.x / aperture).astype(F)
    ph -= ph.mean(dtype=F)
    return ph


def make_vtwins_pupils(config, *, beam_na=0.12, theta_deg=20.0, separation_um=6.0,
                       kind="gaussian", length_um=50.0, profile="gaussian",
                       clip_to_objective=True):
    """pupils.py:54-109: {"left": field, "right": field}, never summed."""
    if (not np.isfinite([beam_na, theta_deg, separation_um, length_um]).all()
            or not 0 < beam_na < config.n or abs(theta_deg) >= 180 or length_um <= 0):
        raise ValueError("Invalid Vtwins beam parameters")
    g = pupil_grid(config)
    aperture = g.objrad
    radius = F(np.tan(np.arcsin(F(beam_na) / g.nr)) * g.fl)
    shift = F(np.tan(np.deg2rad(theta_deg / 2))) * g.fl
    rho2 = g.x * g.x + g.y * g.y
    focus = focus_phase(g)
    ph = tilt_phase(config, g, separation_um, aperture)
    out = {}
    for label, sign in (("left", 1), ("right", -1)):
        u = g.x - sign * shift
        if kind == "gaussian":
            amp = np.exp(-(u * u + g.y * g.y) / (radius * radius)) * (rho2 < aperture * aperture)
        elif kind == "bessel":
            width = F(1.4e-10 * np.tan(np.arcsin(F(config.NA) / g.nr)) / (g.fl * 1e-6 * length_um))
            amp = ring_amplitude(np.sqrt(u * u + g.y * g.y) - radius, width, profile)
            if clip_to_objective:
                amp = amp * (rho2 < aperture * aperture)
        else:
            raise ValueError("Vtwins kind must be gaussian or bessel")
        out[label] = np.asarray(amp * focus * np.exp(np.complex64(1j) * g.k * sign * ph),
                                np.complex64)
    return out
