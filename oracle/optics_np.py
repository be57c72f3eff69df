"""NumPy restatement of the MaoMi optical stage (test infrastructure only).

Citations are to the reference repository (TomaKakuei/MaoMi 0.8.0b1).  The
restatement keeps the reference's operation order, float32 cast points and
shifted FFT sequence so that ``dtype=np.complex64`` reproduces the reference
CPU path to float32 rounding, and ``dtype=np.complex128`` gives the FP64 oracle.

Pinning (see tests/test_oracle.py):
  * geometry, edge_apodization, pupil, resize_xy, translate_xy, fresnel,
    v4_interpolate, propagate, assemble: golden vectors written from the
    reference itself (tests/golden/make_golden.py) plus closed forms
    (Parseval, Q semigroup, Gaussian-beam waist, tilt shift, node
    reproduction of v4, Keys kernel reproduction of constants/ramps).
  * plain_stage_masks, plain_phase, plain_beam, plain_stage (the generic
    split-step workload; tests/test_oracle_plain.py): the energy budget of a
    unitary propagation, free propagation in closed form with zero screens,
    covariance under a joint shift of volume and offsets, equality with the
    golden-pinned segment A of propagate at cut = 0, the reference's own pupil
    vectors and the beam's focal depth, and a direct voxel sum for the screens.
"""

from dataclasses import dataclass
from types import SimpleNamespace
import math

import numpy as np
from scipy import fft as sfft

__all__ = [
    "OracleConfig", "excitation_order", "temporal_profile", "axial_profile",
    "gaussian_beam_size", "vasc_size", "center_slice", "geometry", "edge_apodization",
    "zrf", "zidx", "zernike", "generate_ba", "pupil", "cubic", "resize_xy",
    "translate_xy", "fresnel", "propagate", "v4_interpolate", "assemble",
    "fixed_locations", "locations", "field_shapes", "base_field", "plain_stage_masks",
    "resample_matrix", "set_workers", "plain_phase", "plain_beam", "plain_stage", "zstack",
]

_WORKERS = 1


def set_workers(k):
    """Threads used by scipy.fft in every oracle FFT (CPU baseline knob)."""
    global _WORKERS
    _WORKERS = max(1, int(k))


def _rdt(dtype):
    return np.float32 if np.dtype(dtype) == np.complex64 else np.float64


@dataclass
class OracleConfig:
    """Field-for-field copy of the reference OpticalConfig defaults (optics.py:21-43)."""
    volume_um: tuple = (32.0, 32.0, 10.0)
    depth_um: float = 30.0
    vres: float = 1.0
    NA: float = 0.6
    objNA: float = 0.8
    n: float = 1.35
    wavelength_um: float = 0.92
    focal_length_mm: float = 4.5
    ss: int = 2
    psf_um: tuple = (8.0, 8.0, 10.0)
    prop_um: float = 5.0
    tail_um: float = 20.0
    n_diff: float = 0.02
    hemoabs: float = 0.0155194235
    propcrop: bool = True
    scatter_scales_um: tuple = (0.51, 1.56, 4.52, 14.78)
    scatter_weights: tuple = (0.57, 0.29, 0.19, 0.15)
    seed: int = 20260927
    scaling: str = "two-photon"
    temporal_length_um: float = 10.0
    zernike_weights: tuple = ()


# --------------------------------------------------------------------------- excitation
ORDERS = {"one-photon": 1, "two-photon": 2, "three-photon": 3, "temporal-focusing": 2}


def excitation_order(config):
    """excitation.py:8-11."""
    if config.scaling not in ORDERS:
        raise ValueError(f"Unknown excitation scaling: {config.scaling}")
    return ORDERS[config.scaling]


def temporal_profile(size, length_um, dz_um, *, legacy_offset=0.0):
    """excitation.py:14-29 (closed form)."""
    rayleigh = length_um / (dz_um * 2 * np.sqrt(3))
    indices = np.ceil(size / 2) - np.arange(1, size + 1) + legacy_offset * dz_um
    return 1 / np.sqrt(1 + (indices / rayleigh) ** 2)


def axial_profile(config, size, part="focus"):
    """excitation.py:32-49."""
    if config.scaling != "temporal-focusing":
        return np.ones(size, np.float32)
    if part == "focus":
        dz, offset = 1 / config.vres, 0.0
    else:
        pp = config.prop_um * config.vres
        depth = config.psf_um[2] * config.vres
        dz = pp / config.vres
        offset = ((depth * config.vres + 2 * pp) / (pp * pp) if part == "top"
                  else -depth * config.vres / (pp * pp))
    return temporal_profile(size, config.temporal_length_um, dz,
                            legacy_offset=offset).astype(np.float32)


# --------------------------------------------------------------------------- geometry
def gaussian_beam_size(objNA, n, dist, apod=2.0):
    """_optical_math.py:116-118: radius = ceil(tan(asin(objNA/n))*dist*1.5)*apod."""
    radius = np.ceil(np.tan(np.arcsin(objNA / n)) * dist * 1.5) * apod
    return np.array([radius, radius, 0], dtype=np.float32)


def vasc_size(config):
    """optics.py:67-71 (OpticalConfig.parameters' vasc_sz)."""
    vol = np.array(config.volume_um)
    return (gaussian_beam_size(config.objNA, config.n, config.depth_um + config.volume_um[2] / 2)
            + vol + np.array([0, 0, config.depth_um]))


def center_slice(n, m):
    """_propagation.py:9-12."""
    if m > n or (n - m) % 2:
        raise ValueError(f"Incompatible centered grids {n},{m}")
    return slice((n - m) // 2, (n + m) // 2)


def geometry(config, raw_depths):
    """_propagation.py:15-56; ``raw_depths`` are the last-axis lengths of raw A/B/C."""
    excitation_order(config)
    vasc = vasc_size(config)
    r, ss = config.vres, config.ss
    n = int(round((vasc[0] - np.array(config.volume_um)[0]) * r * ss))
    if config.propcrop:
        b3 = gaussian_beam_size(config.objNA, config.n, config.psf_um[2] / 2, 3)
        n2 = int(round(min(max(b3) * r * 2 * ss, n)))
    else:
        n2 = n
    if n % 2 and not n2 % 2:
        n2 -= 1
    fl = np.float32(config.focal_length_mm / 1000)
    d1 = np.float32(max(gaussian_beam_size(config.objNA, config.n, float(fl * 1e6))) / 1e6 / n)
    d2 = np.float32(1e-6 / r / ss)
    za = r * (config.depth_um + config.volume_um[2] / 2 - config.psf_um[2] / 2)
    pc = config.prop_um * r
    a, b, c = raw_depths
    positions = [
        (np.r_[0, np.arange(a) + (za % pc) / pc] if za % pc else np.arange(a + 1))
        * config.prop_um * 1e-6,
        np.arange(b + 1) / r * 1e-6,
        np.r_[np.arange(c) * config.prop_um, config.tail_um] * 1e-6,
    ]
    cut = max(0, int(np.ceil((za - config.tail_um * r) / pc)))
    if cut >= len(positions[0]) - 1:
        raise ValueError("Top tail has no extent")
    focal_shape = np.round(np.array(config.psf_um[:2]) * r * ss).astype(int)
    return SimpleNamespace(
        n=n, n2=n2, fl=fl, d1=d1, d2=d2,
        wvl=np.float32(config.wavelength_um * 1e-6), nr=np.float32(config.n),
        z=positions, cut=cut, focal_shape=focal_shape,
        crop=center_slice(n, n2),
        focal_crop=tuple(center_slice(n2, int(k)) for k in focal_shape),
    )


def edge_apodization(size, dtype=np.complex64):
    """_propagation.py:59-61: separable super-Gaussian exp(-a^16), a = x/(0.47 size)."""
    rdt = _rdt(dtype)
    axis = np.arange(-size / 2, size / 2, dtype=rdt) / rdt(0.47 * size)
    return (np.exp(-axis[:, None] ** 16) * np.exp(-axis[None, :] ** 16)).astype(rdt)


def fixed_locations(config, count_per_axis=65):
    """fixed.py:15-33."""
    if isinstance(count_per_axis, bool) or not isinstance(count_per_axis, int) or count_per_axis < 2:
        raise ValueError("count_per_axis must be an integer >= 2")
    ticks = np.asarray(config.volume_um[:2], float) * config.vres * config.ss
    if (not np.all(np.isfinite(ticks)) or np.any(ticks < count_per_axis - 1)
            or not np.allclose(ticks, np.round(ticks))):
        raise ValueError("integer optical lattice required")
    axes = [np.round(np.linspace(0, int(round(t)), count_per_axis)) / (config.vres * config.ss)
            for t in ticks]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 2)
    if len(np.unique(pts, axis=0)) != count_per_axis ** 2:
        raise ValueError("Grid contains duplicate locations")
    return pts


def locations(volume_um, count_per_axis):
    """optics.py:75-87."""
    x = (np.linspace(0, volume_um[0], count_per_axis) if count_per_axis > 1
         else np.array([volume_um[0] / 2]))
    y = (np.linspace(0, volume_um[1], count_per_axis) if count_per_axis > 1
         else np.array([volume_um[1] / 2]))
    return np.stack(np.meshgrid(x, y, indexing="ij"), -1).reshape(-1, 2)


def field_shapes(config):
    """fixed.py:36-56 (raw A/B/C shapes)."""
    vasc = vasc_size(config)
    r = config.vres
    n = int(round((vasc[0] - np.array(config.volume_um)[0]) * r))
    if config.propcrop:
        b3 = gaussian_beam_size(config.objNA, config.n, config.psf_um[2] / 2, 3)
        n2 = int(round(min(max(b3) * r * 2, n)))
    else:
        n2 = n
    za = int(round((config.depth_um + config.volume_um[2] / 2 - config.psf_um[2] / 2) * r))
    zb = int(round(config.psf_um[2] * r))
    pc = int(round(config.prop_um * r))
    nc = int(round(config.tail_um * r))
    if min(za, zb, pc, nc) <= 0:
        raise ValueError("Optical A/B/C propagation extents must be positive")
    xy = np.round(np.asarray(config.volume_um[:2]) * r).astype(int)
    return [(xy[0] + n, xy[1] + n, int(np.ceil(za / pc))),
            (xy[0] + n2, xy[1] + n2, zb),
            (xy[0] + n, xy[1] + n, int(np.ceil(nc / pc)))]


# --------------------------------------------------------------------------- pupil
def zrf(n, m, r):
    """_optical_math.py:137-147: radial Zernike polynomial, accumulated in r's dtype."""
    out = np.zeros_like(r)
    for s in range(int((n - m) / 2) + 1):
        num = ((-1) ** s) * math.factorial(int(n - s))
        den = (math.factorial(int(s)) * math.factorial(int((n + m) / 2 - s))
               * math.factorial(int((n - m) / 2 - s)))
        out += (num / den) * (r ** (n - 2 * s))
    return out


def zidx(i, j):
    """_optical_math.py:150-157: Noll index -> (n, m)."""
    idx = math.ceil(math.sqrt(0.25 + 2 * i) - 1.5)
    if j == 1:
        return idx
    return ((idx + 1) % 2) * 2 * math.floor((i - (idx + 1) * idx / 2) / 2) + (idx % 2) * (
        2 * math.ceil((i - (idx + 1) * idx / 2) / 2) - 1)


def zernike(i, x, y):
    """_optical_math.py:160-173: Noll-normalised Zernike Z_i(x, y)."""
    theta = np.arctan2(y, x)
    r = np.sqrt(x ** 2 + y ** 2)
    n, m = zidx(i, 1), zidx(i, 2)
    if m == 0:
        return (math.sqrt(n + 1) * zrf(n, 0, r)).astype(x.dtype)
    trig = np.cos(m * theta) if i % 2 == 0 else np.sin(m * theta)
    return (math.sqrt(2 * (n + 1)) * zrf(n, m, r) * trig).astype(x.dtype)


def _grid_pitch(config, N):
    """Pupil-plane pitch D1 as generate_ba computes it (_optical_math.py:242-247)."""
    fl = float(config.focal_length_mm / 1000.0)
    beam = gaussian_beam_size(config.objNA, config.n, fl * 1e6)
    return float(max(beam) / 1e6 / min(N))


def generate_ba(config, dtype=np.complex64, grid=None, pitch=None):
    """_optical_math.py:232-296 (gaussian profile + optional Zernike screen).

    ``grid`` overrides the sample count (the generic plain-stage workload sets
    it to its field side); by default it follows the vasc/vol geometry.
    """
    rdt = _rdt(dtype)
    fl = float(config.focal_length_mm / 1000.0)
    if grid is None:
        vasc = vasc_size(config)
        D2 = float(1e-6 * (1.0 / config.vres) / config.ss)
        N = np.float32(1e-6 * (vasc[:2] - np.array(config.volume_um)[:2]) / D2)
    else:
        N = np.array([grid, grid], np.float32)
    D1 = _grid_pitch(config, N) if pitch is None else float(pitch)
    nre = float(config.n)
    rad = float(np.tan(np.arcsin(config.NA / nre)) * fl)
    objrad = float(np.tan(np.arcsin(config.objNA / nre)) * fl)
    k = 2 * nre * np.pi / float(config.wavelength_um * 1e-6)
    xr = np.arange(-N[0] / 2, N[0] / 2, dtype=rdt) * rdt(D1)
    yr = np.arange(-N[1] / 2, N[1] / 2, dtype=rdt) * rdt(D1)
    Y, X = np.meshgrid(yr, xr, indexing="ij")
    rho2 = X ** 2 + Y ** 2
    U = np.exp(-(X ** 2 + Y ** 2) / rdt(rad ** 2))
    U = U * (rho2 < rdt(objrad ** 2))
    theta = rdt(-k / (2 * fl)) * rho2
    U = U * (np.cos(theta) + 1j * np.sin(theta)).astype(dtype)
    w = np.asarray(config.zernike_weights, dtype=float).ravel()
    if len(w) and np.any(w != 0):
        abb = w * config.wavelength_um * 1e-6
        Xn, Yn = X / rdt(objrad), Y / rdt(objrad)
        phase = np.zeros_like(Xn)
        for i in np.nonzero(abb)[0]:
            phase += rdt(abb[i]) * zernike(i + 1, Xn, Yn)
        phase = phase - phase.mean(dtype=rdt)
        arg = rdt(k) * phase
        U = U * (np.cos(arg) + 1j * np.sin(arg)).astype(dtype)
    return U.astype(dtype)


def pupil(config, *, custom_phase=None, phase_units="radians", phase_aperture="objective",
          dtype=np.complex64):
    """optics.py:90-117: generate_ba plus an optional custom phase (nearest sampling)."""
    if custom_phase is not None:
        config = SimpleNamespace(**{**vars(config), "zernike_weights": ()})
    u = generate_ba(config, dtype)
    if custom_phase is None:
        return u
    if phase_units not in ("radians", "waves") or phase_aperture not in ("objective", "excitation"):
        raise ValueError("Explicit phase_units / phase_aperture required")
    rdt = _rdt(dtype)
    h, w = u.shape
    fl = config.focal_length_mm / 1000
    d1 = max(gaussian_beam_size(config.objNA, config.n, fl * 1e6)) / 1e6 / min(h, w)
    na = config.objNA if phase_aperture == "objective" else config.NA
    rad = np.tan(np.arcsin(na / config.n)) * fl
    yy = np.arange(-h / 2, h / 2, dtype=rdt) * rdt(d1)
    xx = np.arange(-w / 2, w / 2, dtype=rdt) * rdt(d1)
    y, x = np.meshgrid(yy, xx, indexing="ij")
    gx, gy = x / rdt(rad), y / rdt(rad)
    ph = np.asarray(custom_phase, dtype=rdt)
    H, W = ph.shape
    # grid_sample(mode=nearest, align_corners=True, padding zeros)
    ix = np.rint((gx + 1) / 2 * rdt(W - 1))
    iy = np.rint((gy + 1) / 2 * rdt(H - 1))
    ok = (ix >= 0) & (ix <= W - 1) & (iy >= 0) & (iy <= H - 1)
    sampled = np.zeros_like(gx)
    sampled[ok] = ph[iy[ok].astype(int), ix[ok].astype(int)]
    if phase_units == "waves":
        sampled = sampled * rdt(2 * np.pi)
    return (u * (np.cos(sampled) + 1j * np.sin(sampled))).astype(dtype)


# --------------------------------------------------------------------------- resampling
def cubic(x):
    """numerics.py:8-12: Keys kernel a = -0.5."""
    a = np.abs(x)
    return ((1.5 * a - 2.5) * a * a + 1) * (a <= 1) + (((-0.5 * a + 2.5) * a - 4) * a + 2) * (
        (a > 1) & (a < 2))


def _resize_plan(n, scale):
    m = int(np.ceil(n * scale))
    u = (np.arange(m) + 0.5) / scale - 0.5
    width = 4 / min(scale, 1)
    left = np.floor(u - width / 2).astype(int)
    idx = left[:, None] + np.arange(int(np.ceil(width)) + 2)[None, :]
    weights = cubic((u[:, None] - idx) * min(scale, 1)) * min(scale, 1)
    weights /= weights.sum(1, keepdims=True)
    idx %= 2 * n
    idx = np.where(idx < n, idx, 2 * n - 1 - idx)
    return idx, weights


def resize_xy(a, scale):
    """numerics.py:15-39: MATLAB imresize bicubic (antialias, symmetric extension), X then Y."""
    out = np.asarray(a)
    rdt = out.real.dtype.type
    for axis in (0, 1):
        idx, w = _resize_plan(out.shape[axis], scale)
        moved = np.moveaxis(out, axis, 0)
        gathered = moved[idx]
        wt = w.astype(rdt).reshape(w.shape + (1,) * (moved.ndim - 1))
        out = np.moveaxis((gathered * wt).sum(1), 0, axis)
    return out


def translate_xy(a, shift):
    """numerics.py:79-90: imtranslate bilinear, zero fill: out[i,j] = a(i - s, j - s)."""
    a = np.asarray(a)
    h, w = a.shape[:2]
    f = shift - np.floor(shift)
    k = int(np.floor(shift))
    out = np.zeros_like(a)
    # out[i] = (1-f) a[i-k] + f a[i-k-1] along each axis (zero outside).
    def shifted(arr, axis, s):
        res = np.zeros_like(arr)
        n = arr.shape[axis]
        src = [slice(None)] * arr.ndim
        dst = [slice(None)] * arr.ndim
        if s >= n or -s >= n:
            return res
        if s >= 0:
            src[axis], dst[axis] = slice(0, n - s), slice(s, n)
        else:
            src[axis], dst[axis] = slice(-s, n), slice(0, n + s)
        res[tuple(dst)] = arr[tuple(src)]
        return res
    rdt = a.real.dtype.type
    g, fr = rdt(1 - f), rdt(f)
    tmp = g * shifted(a, 0, k) + fr * shifted(a, 0, k + 1)
    out = g * shifted(tmp, 1, k) + fr * shifted(tmp, 1, k + 1)
    return out.astype(a.dtype)


def resample_matrix(n, ss):
    """native.py:27-48: translate(ss/2-0.5) then imresize(1/ss) as one dense (m, n) matrix."""
    scale = 1 / ss
    shift = ss / 2 - 0.5
    idx, weights = _resize_plan(n, scale)
    m = idx.shape[0]
    mat = np.zeros((m, n))
    for i in range(m):
        for j, w in zip(idx[i], weights[i]):
            a = int(np.floor(j - shift))
            f = j - shift - a
            if 0 <= a < n:
                mat[i, a] += w * (1 - f)
            if 0 <= a + 1 < n:
                mat[i, a + 1] += w * f
    return mat


# --------------------------------------------------------------------------- propagation
def _fft2(x):
    return sfft.fft2(x, axes=(-2, -1), workers=_WORKERS)


def _ifft2(x):
    return sfft.ifft2(x, axes=(-2, -1), workers=_WORKERS)


def fresnel(u, wavelength, dx, z, phi, nidx, *, crop=None, save_from=0, return_all=False):
    """numerics.py:156-210: scaled multi-screen Fresnel with the shifted FFT sequence.

    u (..., X, Y); phi (..., X, Y, S) with S = len(z) or len(z) - 1.  Arithmetic
    follows u's precision (complex64 mirrors the reference float32 cast points).
    """
    dx, z = np.asarray(dx, float).ravel(), np.asarray(z, float).ravel()
    if len(z) < 2 or len(dx) != len(z) or np.any(np.diff(z) <= 0):
        raise ValueError("Need increasing z and one dx per plane")
    if phi.shape[-1] not in (len(z), len(z) - 1):
        raise ValueError("Wrong number of phase screens")
    cdt = u.dtype
    rdt = u.real.dtype.type
    h, w = u.shape[-2:]
    x = np.arange(-h / 2, h / 2, dtype=rdt)[:, None]
    y = np.arange(-w / 2, w / 2, dtype=rdt)[None, :]
    lam = float(wavelength) / float(nidx)
    k = 2 * np.pi / lam
    dz = np.diff(z)
    sc = dx[1:] / dx[:-1]

    def cexp(theta):
        return (np.cos(theta) + 1j * np.sin(theta)).astype(cdt)

    # Python evaluates 1j*k*T*(1-sc)/(2dz) left to right on a complex64 tensor; the final
    # division by a real scalar multiplies by its rounded reciprocal.
    entry = rdt(k) * ((x * rdt(dx[0])) ** 2 + (y * rdt(dx[0])) ** 2)
    entry = (entry * rdt(1 - sc[0])) * (rdt(1) / rdt(2 * dz[0]))
    u = u * cexp(entry) * phi[..., 0]
    saved = []

    def keep(i, field):
        if return_all and i >= save_from:
            saved.append(field[..., crop[0], crop[1]].copy() if crop else field.copy())

    keep(0, u)
    cache = {}
    for i in range(len(dz)):
        key = (float(dz[i]), float(dx[i]), float(sc[i]))
        if key not in cache:
            v = (x / rdt(h * dx[i])) ** 2 + (y / rdt(w * dx[i])) ** 2
            arg = rdt(-np.pi * lam * dz[i]) * v
            if sc[i] != 1:
                # complex / real scalar: multiply by the rounded reciprocal (c10 complex division)
                arg = arg * (rdt(1) / rdt(sc[i]))
            cache[key] = cexp(arg)
        uu = u if sc[i] == 1 else (u * (rdt(1) / rdt(sc[i]))).astype(cdt)
        f = sfft.fftshift(_fft2(sfft.fftshift(uu, axes=(-2, -1))), axes=(-2, -1))
        u = sfft.ifftshift(_ifft2(sfft.ifftshift(cache[key] * f, axes=(-2, -1))),
                           axes=(-2, -1)).astype(cdt)
        if i + 1 < phi.shape[-1]:
            u = u * phi[..., i + 1]
        keep(i + 1, u)
    exitc = rdt(k / 2 * (sc[-1] - 1) / (sc[-1] * dz[-1])) * (
        (x * rdt(dx[-1])) ** 2 + (y * rdt(dx[-1])) ** 2)
    out = (u * cexp(exitc)).astype(cdt)
    return (out, np.stack(saved, axis=-1)) if return_all else out


def base_field(u, wvl, d1, d2, dz, nr, sg, norm="reference"):
    """batched.py:43-51 / optics.py:223-231: one scaled step d1 -> d2 then unit L2 norm.

    norm="reference" mirrors torch.linalg.vector_norm on complex64 in the
    reference: |b| squared and summed strictly sequentially in the field's real
    precision (this under-estimates the norm by ~4e-6 at n=92).  norm="exact"
    takes the norm in FP64.
    """
    base = fresnel(u, wvl, [d1, d2], [0, dz], np.stack((sg, sg), -1), nr)
    rdt = base.real.dtype.type
    if norm == "reference":
        a = np.abs(base).ravel()
        nrm = np.sqrt(np.cumsum(a * a, dtype=rdt)[-1])
    elif norm == "exact":
        nrm = rdt(np.sqrt(np.sum(np.abs(base.astype(np.complex128)) ** 2)))
    else:
        raise ValueError("norm must be reference or exact")
    return (base * (rdt(1) / rdt(nrm))).astype(u.dtype)


def _power(field, order):
    p = np.abs(field)
    return p ** (2 * order)


def propagate(config, raw_fields, points_um, *, uin=None, dtype=np.complex64, save_full=False,
              norm="exact"):
    """optics.py:200-381 (phase_mode='full'): serial per-site A/B/C split-step, tails, v4.

    The base beam's unit-norm scaling uses an FP64 norm by default (see base_field).
    """
    cdt = np.dtype(dtype)
    rdt = _rdt(cdt)
    r, ss = config.vres, config.ss
    g = geometry(config, [a.shape[-1] for a in raw_fields])
    n, n2 = g.n, g.n2
    u = pupil(config, dtype=cdt) if uin is None else np.asarray(uin).astype(cdt)
    sg, sg2 = edge_apodization(n, cdt), edge_apodization(n2, cdt)
    dz0 = np.float32(g.fl - (config.depth_um + config.volume_um[2] / 2) * 1e-6)
    base = base_field(u, g.wvl, g.d1, g.d2, dz0, g.nr, sg, norm=norm)
    phase_scale = config.n_diff * 2 * np.pi / (config.wavelength_um * r)
    phase = []
    for a in raw_fields:
        arg = resize_xy(np.asarray(a, dtype=rdt), ss) * rdt(phase_scale)
        phase.append((np.cos(arg) + 1j * np.sin(arg)).astype(cdt))
    za_pos, zb_pos, zc_pos = g.z
    cut, c, fc = g.cut, g.crop, g.focal_crop
    order = excitation_order(config)
    outputs, tails, weights, profiles = [], [[], []], [[], []], [[], []]
    pts = np.asarray(points_um, float)
    if np.any(pts < 0) or np.any(pts > np.array(config.volume_um[:2])):
        raise ValueError("PSF coordinates outside FOV")
    for xy in pts:
        offset = xy * r * ss
        if not np.allclose(offset, np.round(offset)):
            raise ValueError("Locations must land on the supersampled grid")
        x, y = np.round(offset).astype(int)
        pa = phase[0][x:x + n, y:y + n] * sg[..., None]
        pb = phase[1][x:x + n2, y:y + n2] * sg2[..., None]
        pcf = phase[2][x:x + n, y:y + n] * sg[..., None]
        ua, top = fresnel(base, g.wvl, np.full(len(za_pos), g.d2), za_pos, pa, g.nr,
                          return_all=True, crop=None if save_full else (c, c),
                          save_from=0 if save_full else cut)
        ub, focus = fresnel(ua[c, c], g.wvl, np.full(len(zb_pos), g.d2), zb_pos, pb, g.nr,
                            return_all=True, crop=None if save_full else fc)
        pad = (n - n2) // 2
        padded = np.pad(ub, ((pad, pad), (pad, pad)))
        _, bottom = fresnel(padded, g.wvl, np.full(len(zc_pos), g.d2), zc_pos, pcf, g.nr,
                            return_all=True, crop=None if save_full else (c, c))
        if save_full:
            top, focus, bottom = top[c, c, cut:], focus[fc[0], fc[1]], bottom[c, c]
        intensity = rdt(ss ** 2) * _power(focus[..., :-1], order)
        intensity = intensity * axial_profile(config, intensity.shape[-1]).astype(rdt)
        final = (rdt(ss ** 2) * resize_xy(translate_xy(intensity, ss / 2 - 0.5), 1 / ss)
                 * rdt(r * (1e6 * g.wvl) ** 1.5) / rdt(np.pi * g.nr))
        outputs.append(final.astype(rdt))
        for j, field in enumerate((top, bottom)):
            inten = rdt(ss ** 2) * _power(field, order)
            inten = inten * axial_profile(config, field.shape[-1],
                                          "top" if j == 0 else "bottom").astype(rdt)
            profiles[j].append(inten.sum((0, 1)))
            inten[..., 0] *= rdt(0.5)
            inten[..., -1] *= rdt(0.5)
            weights[j].append(float(inten.sum()) * config.prop_um)
            tails[j].append(rdt(ss ** 2) * resize_xy(translate_xy(inten.sum(-1), ss / 2 - 0.5),
                                                     1 / ss))
    psfs = np.stack(outputs)
    energies = psfs.sum((1, 2, 3), dtype=np.float64)
    xyshape = np.round(np.array(config.volume_um[:2]) * r).astype(int)
    q = np.stack(np.meshgrid(np.arange(xyshape[0]) + 0.5, np.arange(xyshape[1]) + 0.5,
                             indexing="ij"), -1)
    mask = v4_interpolate(pts * r, energies, q)
    mean = float(mask.mean())
    result = {
        "psf_3d": psfs, "psf_raw": psfs.mean(0), "psf": np.abs(psfs).mean(0),
        "mask_raw": mask, "mask_mean": mean, "mask": mask / mean if mean > 0 else mask,
        "points_um": pts, "energies": energies,
        "negative_resampling_mass_fraction": float(
            -np.minimum(psfs, 0).sum(dtype=np.float64) / np.abs(psfs).sum(dtype=np.float64)),
        "z_positions": g.z, "cut": cut,
    }
    for j, name in enumerate(("psfT", "psfB")):
        profile = np.stack(profiles[j]).sum(0)
        actual_z = (za_pos[cut:] - za_pos[cut]) if j == 0 else zc_pos
        out_z = np.arange(round(actual_z[-1] * 1e6 * r) + 1) / r
        interp = np.interp(out_z, actual_z * 1e6, profile)
        tail_mask = v4_interpolate(pts * r, weights[j], q)
        tail_mean = float(tail_mask.mean())
        result[name] = {
            "convmask": np.mean(tails[j], 0), "weight": float(np.mean(weights[j])),
            "psfZ": interp / interp.mean() if interp.mean() > 0 else interp,
            "mask_raw": tail_mask, "mask_mean": tail_mean,
            "mask": tail_mask / tail_mean if tail_mean > 0 else tail_mask,
            "weights_per_point": np.asarray(weights[j]),
            "profiles_per_point": np.stack(profiles[j]),
            "convmask_per_point": np.stack(tails[j]),
        }
    return result


# --------------------------------------------------------------------------- assembly
def _green(x, y):
    d = np.sqrt(((x[:, None, :] - y[None, :, :]) ** 2).sum(-1))
    with np.errstate(divide="ignore", invalid="ignore"):
        g = d * d * (np.log(d) - 1)
    g[d == 0] = 0
    return g


def v4_interpolate(points, values, queries, chunk=16384):
    """numerics.py:93-153 (CPU evaluator): MATLAB griddata 'v4' biharmonic spline in FP64."""
    from scipy.spatial.distance import cdist
    if not isinstance(chunk, (int, np.integer)) or chunk < 1:
        raise ValueError("chunk must be a positive integer")
    p, v, q = np.asarray(points, float), np.asarray(values, float), np.asarray(queries, float)
    if p.ndim != 2 or p.shape[1] != 2 or not len(p):
        raise ValueError("points must be a nonempty (N, 2) array")
    if v.ndim not in (1, 2) or len(v) != len(p) or (v.ndim == 2 and v.shape[1] == 0):
        raise ValueError("values must have shape (N,) or (N, C) with C > 0")
    if q.ndim < 1 or q.shape[-1] != 2:
        raise ValueError("queries must have shape (..., 2)")
    shape = q.shape[:-1] + v.shape[1:]
    if len(p) == 1:
        return np.broadcast_to(v[0], shape).copy()
    flat = q.reshape(-1, 2)
    if not len(flat):
        return np.empty(shape, dtype=np.float64)
    coef = np.linalg.solve(_green(p, p), v)
    result = np.empty((len(flat),) + v.shape[1:], dtype=np.float64)
    chunk = min(chunk, max(1, 8_388_608 // len(p)))
    for start in range(0, len(flat), chunk):
        d2 = cdist(flat[start:start + chunk], p, "sqeuclidean")
        with np.errstate(divide="ignore", invalid="ignore"):
            basis = np.log(d2)
            basis *= 0.5
            basis -= 1
            basis *= d2
        basis[d2 == 0] = 0
        result[start:start + chunk] = basis @ coef
    return result.reshape(shape)


def assemble(config, out, shape, *, z_positions=None, cut=None):
    """fixed.py:119-190 (mask_mode='matlab_v4')."""
    points, psfs = out["points_um"], out["psfs"]
    energies = psfs.sum((1, 2, 3), dtype=np.float64)
    values = [energies]
    tails = bool(out["tails_computed"])
    if tails:
        values.extend(out[k]["weight"] for k in ("top", "bottom"))
    values = np.stack(values, axis=1)
    axes = [(np.arange(k) + 0.5) for k in shape]
    q = np.stack(np.meshgrid(*axes, indexing="ij"), -1)
    masks = v4_interpolate(points * config.vres, values, q)
    weights = np.full(len(points), 1 / len(points))
    mask = masks[..., 0]
    mean = float(mask.mean())
    if not np.isfinite(mean) or mean <= 0:
        raise ValueError("Nonpositive or invalid mean excitation energy")
    result = {
        "psf": (weights @ np.abs(psfs).reshape(len(points), -1)).reshape(psfs.shape[1:]),
        "psf_raw": (weights @ psfs.reshape(len(points), -1)).reshape(psfs.shape[1:]),
        "psf_3d": psfs, "points_um": points, "energies": energies,
        "mask_raw": mask, "mask": mask / mean, "mask_mean": mean,
        "tails_computed": tails, "unique_optical_paths": len(points),
        "negative_resampling_mass_fraction": float(
            -np.minimum(psfs, 0).sum(dtype=np.float64)
            / max(np.abs(psfs).sum(dtype=np.float64), 1e-30)),
    }
    if tails:
        for j, (native, name) in enumerate([("top", "psfT"), ("bottom", "psfB")]):
            data = out[native]
            profile = data["profile"].sum(0, dtype=np.float64)
            actual_z = z_positions[0][cut:] if j == 0 else z_positions[2]
            actual_z = actual_z - actual_z[0]
            axial = np.arange(round(actual_z[-1] * 1e6 * config.vres) + 1) / config.vres
            interp = np.interp(axial, actual_z * 1e6, profile)
            tm = masks[..., j + 1]
            tm_mean = float(tm.mean())
            result[name] = {
                "convmask": data["convmask"].mean(0), "weight": float(data["weight"].mean()),
                "psfZ": interp / interp.mean() if interp.mean() > 0 else interp,
                "mask_raw": tm, "mask": tm / tm_mean if tm_mean > 0 else tm,
                "mask_mean": tm_mean, "weights_per_point": data["weight"],
                "profiles_per_point": data["profile"],
            }
    return result


# --------------------------------------------------------------------------- plain stage
def plain_stage_masks(base, phase, offsets, *, wvl, nr, dx, dz, sizes=None, slices=None,
                      edge=True, dtype=np.complex64, order=2, return_planes=False):
    """Generic per-site split-step of the plain optical stage workload.

    Per site: fresnel(base, wvl, const dx, z = arange(S+1) dz, phi = window(phase)
    * edge, nr, return_all=True) -- the same call optics.py:282-292 makes for
    segment A with cut = 0 -- then M = sum_p trap_p |u_p|^(2 order) over the S+1
    saved planes (trap = 1/2 at both ends, optics.py:333-334 / batched.py:217-219).

    base: dict N -> (N, N) complex field or one (N, N) field; phase: (Z, H, W)
    float radians; offsets: (P, 2) window top-left corners; sizes/slices: per
    site N and S (default: base side and Z).
    """
    cdt = np.dtype(dtype)
    rdt = _rdt(cdt)
    offsets = np.asarray(offsets, int)
    P = len(offsets)
    bases = base if isinstance(base, dict) else {base.shape[-1]: base}
    if sizes is None:
        sizes = np.full(P, next(iter(bases)))
    if slices is None:
        slices = np.full(P, phase.shape[0])
    out = []
    planes_out = []
    for i in range(P):
        N, S = int(sizes[i]), int(slices[i])
        b = bases[N] if not isinstance(bases[N], dict) else bases[N][S]
        x, y = offsets[i]
        win = np.asarray(phase[:S, x:x + N, y:y + N], dtype=rdt)
        scr = (np.cos(win) + 1j * np.sin(win)).astype(cdt)
        if edge:
            scr = scr * edge_apodization(N, cdt)[None]
        scr = np.moveaxis(scr, 0, -1)
        z = np.arange(S + 1) * dz
        _, saved = fresnel(b.astype(cdt), wvl, np.full(S + 1, dx), z, scr, nr, return_all=True)
        inten = _power(saved, order).astype(rdt)
        inten[..., 0] *= rdt(0.5)
        inten[..., -1] *= rdt(0.5)
        out.append(inten.sum(-1))
        if return_planes:
            planes_out.append(saved)
    return (out, planes_out) if return_planes else out


def plain_phase(index_volume, n0, k0dz, voxels_per_slice=1, dtype=np.complex64):
    """Screens of the plain stage: phi[z] = k0 dz sum_v (n - n0), accumulated in FP64.

    maomi/optics.py:232 (phase_scale * raw, raw = (n - n0)/n_diff summed over the slab).
    index_volume (X, Y, Zv) -> phase (Z, X, Y).
    """
    v = np.asarray(index_volume)
    X, Y, Zv = v.shape
    acc = (v.astype(np.float64) - n0).reshape(X, Y, Zv // voxels_per_slice, voxels_per_slice)
    acc = acc.sum(-1)
    ph = (k0dz * acc).astype(_rdt(dtype))
    return np.ascontiguousarray(np.moveaxis(ph, -1, 0))


def plain_beam(cfg, zernike_weights=(), dtype=np.complex64):
    """Pupil (generate_ba on an N-sample grid) carried to the entry plane, unit norm.

    _optical_math.py:232-296 with N = cfg.n and D1 = beam_size(fl)/N; numerics.py:156-210
    (one scaled step d1 -> dx over fl - focus); batched.py:51 (unit L2 norm, FP64 here).
    """
    N = int(cfg.n)
    fl = float(cfg.focal_length_mm / 1000.0)
    conf = OracleConfig(NA=cfg.NA, objNA=cfg.objNA, n=cfg.n0, wavelength_um=cfg.wavelength_um,
                        focal_length_mm=cfg.focal_length_mm,
                        zernike_weights=tuple(zernike_weights))
    d1 = _grid_pitch(conf, np.array([N, N], np.float32))
    u = generate_ba(conf, dtype, grid=N, pitch=d1)
    sg = edge_apodization(N, dtype)
    wvl, nr = np.float32(cfg.wavelength_um * 1e-6), np.float32(cfg.n0)
    dx = np.float32(cfg.dx_um * 1e-6)
    dzf = np.float32(fl - cfg.focus * 1e-6)
    return base_field(u, wvl, d1, dx, dzf, nr, sg, norm="exact"), u


def plain_stage(cfg, index_volume, offsets, zernike_weights=(), dtype=np.complex64,
                sites=None, with_masks=True):
    """The whole plain optical stage on the CPU: beam, screens, per-site masks, v4 FOV mask."""
    base, u = plain_beam(cfg, zernike_weights, dtype)
    k0dz = 2 * np.pi / cfg.wavelength_um * cfg.dz_um / cfg.voxels_per_slice
    phase = plain_phase(index_volume, cfg.n0, k0dz, cfg.voxels_per_slice, dtype)
    offs = np.asarray(offsets, int)
    idx = np.arange(len(offs)) if sites is None else np.asarray(sites)
    masks = plain_stage_masks(base, phase, offs[idx], wvl=np.float32(cfg.wavelength_um * 1e-6),
                              nr=np.float32(cfg.n0), dx=np.float32(cfg.dx_um * 1e-6),
                              dz=cfg.dz_um * 1e-6, dtype=dtype)
    masks = np.stack(masks)
    energies = masks.sum((1, 2), dtype=np.float64)
    out = dict(masks=masks, energies=energies, base=base, pupil=u, phase=phase)
    if with_masks and sites is None:
        q = np.stack(np.meshgrid(np.arange(cfg.fov) + 0.5, np.arange(cfg.fov) + 0.5,
                                 indexing="ij"), -1)
        m = v4_interpolate(offs.astype(np.float64), energies, q)
        out.update(mask_raw=m, mask_mean=float(m.mean()), mask=m / m.mean())
    return out


def zstack(volume, psf, mask=None, collection_mask=None):
    """scanning.py:8-53 (matlab_convn): convn(volume * masks, psf, 'same') with the MATLAB crop
    floor(k/2), zero extension, FP64.  Pinned by a direct sum in tests/test_oracle.py."""
    from scipy.signal import fftconvolve
    v = np.asarray(volume, np.float64)
    for m in (mask, collection_mask):
        if m is not None:
            v = v * np.asarray(m, np.float64)[..., None]
    p = np.asarray(psf, np.float64)
    full = fftconvolve(v, p, mode="full")
    return full[tuple(slice(b // 2, b // 2 + a) for a, b in zip(v.shape, p.shape))]
