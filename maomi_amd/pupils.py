"""Alternative input pupils: make_pupil and make_vtwins_pupils (maomi/pupils.py:7-122).

The fields are the `input_pupil` of run_fixed (a mapping of them runs one channel per beam,
fixed.py:239-287).  The host computes what is O(n): the float32 scalars of each expression in
the reference's own order, the grid's coordinate vectors (pupils.py:37-51) and, for the pair
beams, the mean-removed tilt phase per column (pupils.py:99-105; its float32 mean is NumPy's
over the whole grid).  maomi_pupil_field computes every pixel on the GPU.  Fields come back as
complex64 tensors on `device`; the reference returns NumPy arrays of the same values.
"""

from types import SimpleNamespace

import numpy as np
import torch

from . import _native as nat
from ._optical_math import gaussian_beam_size

F = np.float32
_PROFILES = {"gaussian": 0, "uniform": 1, "posax": 2, "negax": 3}


def _grid(config):
    """pupils.py:37-51: float32 pitch, radii, wavenumber and coordinate vectors."""
    v, p = config.parameters()
    fl = F(config.focal_length_mm / 1000)
    d2 = F(1e-6 / config.vres / config.ss)
    n = F(1e-6 * (v.vasc_sz[:2] - v.vol_sz[:2]) / d2)
    d1 = F(max(gaussian_beam_size(p, float(fl * 1e6))) / 1e6 / min(n))
    nr = F(config.n)
    rad = F(np.tan(np.arcsin(F(config.NA) / nr)) * fl)
    objrad = F(np.tan(np.arcsin(F(config.objNA) / nr)) * fl)
    k = F(2 * nr * np.pi / F(config.wavelength_um * 1e-6))
    xs = np.arange(-n[0] / 2, n[0] / 2, dtype=F) * d1
    ys = np.arange(-n[1] / 2, n[1] / 2, dtype=F) * d1
    focus = F((np.complex64(-1j) * k / (2 * fl)).imag)  # exp(focus * rho2 i)
    return SimpleNamespace(xs=xs, ys=ys, rad=rad, objrad=objrad, k=k, fl=fl, nr=nr, focus=focus)


def _ring(params, width, profile):
    """The ring profile's own width (pupils.py:112-122) into params."""
    if profile not in _PROFILES:
        raise ValueError("Unknown Bessel profile")
    params.profile = _PROFILES[profile]
    if profile == "gaussian":
        params.width = float(width / (2 * np.sqrt(F(np.log(2)))))
    elif profile == "uniform":
        params.width = float(width / 2)
    else:
        w = width / 1.133
        params.width, params.offset = float(w), float(w / np.sqrt(F(2)))


def _field(g, params, device, cols=None):
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("maomi_amd.pupils builds the fields on the GPU; pass a cuda device")
    xs = torch.as_tensor(g.xs, device=dev)
    ys = torch.as_tensor(g.ys, device=dev)
    c = None if cols is None else torch.as_tensor(np.ascontiguousarray(cols, F), device=dev)
    out = torch.empty((len(g.ys), len(g.xs)), dtype=torch.complex64, device=dev)
    nat.check(nat.lib().maomi_pupil_field(nat.ptr(out), len(g.xs), len(g.ys), nat.ptr(xs),
                                          nat.ptr(ys), nat.ptr(c), params,
                                          nat.stream_handle(dev)), "pupil field")
    return out


def _params(g, **kw):
    p = nat.PupilParams()
    kw.update(obj2=g.objrad * g.objrad, focus=g.focus)
    for k, v in kw.items():
        setattr(p, k, int(v) if isinstance(v, (int, np.integer)) else float(v))
    return p


def make_pupil(config, *, kind="gaussian", length_um=60.0, profile="gaussian",
               cylscale=(1.0, 1.0), device="cuda"):
    """One input field (pupils.py:7-34): kind gaussian, cylindrical or bessel."""
    g = _grid(config)
    if kind == "gaussian":
        params = _params(g, kind=0, rr=g.rad * g.rad)
    elif kind == "cylindrical":
        scale = np.asarray(cylscale, float)
        if scale.shape != (2,) or not np.isfinite(scale).all() or np.any(scale <= 0):
            raise ValueError("Positive two-axis cylindrical scale required")
        params = _params(g, kind=0, rr=g.rad * g.rad, cylinder=1, scale_x=float(scale[0]),
                         scale_y=float(scale[1]))
    elif kind == "bessel":
        if not np.isfinite(length_um) or length_um <= 0:
            raise ValueError("Positive Bessel length required")
        width = F(1.4e-10 * np.tan(np.arcsin(F(config.NA) / g.nr)) / (g.fl * 1e-6 * 2 * length_um))
        params = _params(g, kind=1, radius=g.rad)
        _ring(params, width, profile)
    else:
        raise ValueError("Supported single-field kinds: gaussian, bessel, cylindrical")
    field = _field(g, params, device)
    if not bool(torch.isfinite(torch.view_as_real(field)).all()) or not bool((field != 0).any()):
        raise ValueError("Pupil has no finite nonzero energy on this grid")
    return field


def make_vtwins_pupils(config, *, beam_na=0.12, theta_deg=20.0, separation_um=6.0,
                       kind="gaussian", length_um=50.0, profile="gaussian",
                       clip_to_objective=True, device="cuda"):
    """The vTwINS pair (pupils.py:54-109): {"left": field, "right": field}, two separate
    incident fields, never a coherent sum.  clip_to_objective applies to the bessel kind."""
    if (not np.isfinite([beam_na, theta_deg, separation_um, length_um]).all()
            or not 0 < beam_na < config.n or abs(theta_deg) >= 180 or length_um <= 0):
        raise ValueError("Invalid Vtwins beam parameters")
    if kind not in ("gaussian", "bessel"):
        raise ValueError("Vtwins kind must be gaussian or bessel")
    g = _grid(config)
    aperture = g.objrad
    radius = F(np.tan(np.arcsin(F(beam_na) / g.nr)) * g.fl)
    shift = F(np.tan(np.deg2rad(theta_deg / 2))) * g.fl
    # the tilt phase depends on x only; NumPy's float32 mean over the full grid sets its offset
    x2d = np.meshgrid(g.xs, g.ys)[0]
    ph = (F(config.wavelength_um * 1e-6 * separation_um / 8) * F(2) * x2d / aperture).astype(F)
    cols = ph[0] - ph.mean(dtype=F)
    out = {}
    for label, sign in (("left", 1), ("right", -1)):
        tilt = F((np.complex64(1j) * g.k * sign).imag)
        if kind == "gaussian":
            params = _params(g, kind=2, rr=radius * radius, shift=sign * shift, tilt=tilt)
        else:
            width = F(1.4e-10 * np.tan(np.arcsin(F(config.NA) / g.nr)) / (g.fl * 1e-6 * length_um))
            params = _params(g, kind=3, radius=radius, shift=sign * shift, tilt=tilt,
                             clip=int(bool(clip_to_objective)))
            _ring(params, width, profile)
        out[label] = _field(g, params, device, cols)
    return out
