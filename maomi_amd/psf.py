"""2-D PSF data contracts and the callable PSF-provider boundary.

Keeps the reference's public contract (maomi/psf.py:13-203) so that kernels, fields and files
pass between the two packages unchanged: ``SampledPSF``, ``PSFProvider``, ``PSFField`` (the
same validation and error texts, the same NPZ keys and values), ``resample_psf`` and
``pupil_to_intensity``.  Coordinates are specimen-plane micrometres, arrays are XY, kernel
origins sit at shape // 2.

The containers are host FP64 arrays, frozen and read-only, as in the reference.  The arithmetic
of ``resample_psf`` (bilinear sampling) and ``pupil_to_intensity`` (pupil, FFT, power) runs on
the GPU in FP64 (``maomi_resample_linear_f64``, ``maomi_pupil_embed_f64``, an FP64 FFT,
``maomi_power_shift_f64``); there is no CPU path.  ``maomi_amd.providers.OpticalStagePSF`` is
the GPU optical stage as a provider.
"""

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Protocol

import numpy as np

_FIELD_AXES = "field_X,field_Y,kernel_X,kernel_Y"
_FIELD_INTERP = "source_bilinear_edge"
_COORD_MSG = "PSF field coordinates must be finite and strictly increasing"


def _frozen(a):
    a.setflags(write=False)
    return a


def _json_roundtrip(meta):
    """Provenance as plain JSON (no NaN), which is what the NPZ files store."""
    return json.loads(json.dumps(meta, allow_nan=False))


def _checked_kernel(intensity):
    a = np.array(intensity, dtype=np.float64, copy=True)
    if a.ndim != 2 or min(a.shape) < 2 or not np.isfinite(a).all() or (a < 0).any():
        raise ValueError("PSF must be a finite, nonnegative 2-D intensity array")
    if not np.isclose(a.sum(), 1.0, rtol=1e-10, atol=1e-12):
        raise ValueError("PSF must have unit sum; use SampledPSF.normalized explicitly")
    return a


def _checked_pitch(pixel_size_um):
    pitch = tuple(float(v) for v in pixel_size_um)
    if len(pitch) != 2 or not np.isfinite(pitch).all() or min(pitch) <= 0:
        raise ValueError("Specify two positive specimen-plane pixel pitches in micrometres")
    return pitch


def _checked_nodes(values):
    c = np.array(values, dtype=float, copy=True)
    if c.ndim != 1 or c.size == 0 or not np.isfinite(c).all() or (np.diff(c) <= 0).any():
        raise ValueError(_COORD_MSG)
    return c


@dataclass(frozen=True)
class SampledPSF:
    """Unit-sum redistribution kernel centred at shape // 2 (a shift-invariant, incoherent,
    monochromatic 2-D detection PSF).  The pixel pitch is in specimen-plane micrometres;
    throughput and illumination belong to the acquisition model, not to this kernel."""

    intensity: np.ndarray
    pixel_size_um: tuple
    wavelength_nm: float
    provenance: dict = field(default_factory=dict)

    def __post_init__(self):
        a = _checked_kernel(self.intensity)
        pitch = _checked_pitch(self.pixel_size_um)
        if not np.isfinite(self.wavelength_nm) or self.wavelength_nm <= 0:
            raise ValueError("wavelength_nm must be positive and finite")
        meta = _json_roundtrip(self.provenance)
        object.__setattr__(self, "intensity", _frozen(a))
        object.__setattr__(self, "pixel_size_um", pitch)
        object.__setattr__(self, "provenance", meta)

    @classmethod
    def normalized(cls, intensity, pixel_size_um, wavelength_nm, provenance=None):
        a = np.asarray(intensity, dtype=np.float64)
        total = a.sum()
        if not np.isfinite(a).all() or (a < 0).any() or not np.isfinite(total) or total <= 0:
            raise ValueError("Cannot normalize an invalid or zero-energy intensity PSF")
        return cls(a / total, pixel_size_um, wavelength_nm, provenance or {})

    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, intensity=self.intensity, pixel_size_um=self.pixel_size_um,
                            wavelength_nm=self.wavelength_nm, axes="XY",
                            normalization="unit_sum", provenance=json.dumps(self.provenance))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as d:
            if str(d["axes"]) != "XY" or str(d["normalization"]) != "unit_sum":
                raise ValueError("Unsupported PSF axes or normalization")
            return cls(d["intensity"], tuple(d["pixel_size_um"]), float(d["wavelength_nm"]),
                       json.loads(str(d["provenance"])))


class PSFProvider(Protocol):
    """Any callable (x_um, y_um) -> SampledPSF at specimen coordinates; a function or a
    callable object is enough, with no registration or inheritance.  A provider may also have
    ``prefetch(x_um, y_um)``, which PSFField.from_provider calls once with the whole grid
    before sampling it node by node (the GPU stage evaluates every node in one launch)."""

    def __call__(self, x_um: float, y_um: float) -> SampledPSF: ...


@dataclass(frozen=True)
class PSFField:
    """Unit-sum kernels on a rectilinear grid of source positions, axes field_X, field_Y,
    kernel_X, kernel_Y, with one sampling shared by every kernel.  Bilinear interpolation at
    emitting positions, nearest-edge extension outside the grid, a singleton axis constant.
    An incoherent, monochromatic 2-D operator, not a complex-field model."""

    intensity: np.ndarray
    x_um: np.ndarray
    y_um: np.ndarray
    pixel_size_um: tuple
    wavelength_nm: float
    provenance: dict = field(default_factory=dict)

    def __post_init__(self):
        a = np.array(self.intensity, dtype=np.float64, copy=True)
        x, y = _checked_nodes(self.x_um), _checked_nodes(self.y_um)
        if a.ndim != 4 or a.shape[:2] != (len(x), len(y)):
            raise ValueError("PSF field axes must be field_X, field_Y, kernel_X, kernel_Y")
        # every kernel must itself be a valid SampledPSF of the field's sampling
        kernels = a.reshape(-1, *a.shape[2:])
        first = SampledPSF(kernels[0], self.pixel_size_um, self.wavelength_nm, self.provenance)
        for kernel in kernels[1:]:
            _checked_kernel(kernel)
        object.__setattr__(self, "intensity", _frozen(a))
        object.__setattr__(self, "x_um", _frozen(x))
        object.__setattr__(self, "y_um", _frozen(y))
        object.__setattr__(self, "pixel_size_um", first.pixel_size_um)
        object.__setattr__(self, "provenance", first.provenance)

    @classmethod
    def from_provider(cls, provider, x_um, y_um):
        """Sample a provider once per grid node; kernels must share shape, pitch and wavelength
        (never silently resampled)."""
        x = _checked_nodes(np.asarray(x_um, dtype=float))
        y = _checked_nodes(np.asarray(y_um, dtype=float))
        prefetch = getattr(provider, "prefetch", None)
        if callable(prefetch):
            prefetch(x, y)
        grid = np.empty((len(x), len(y)), dtype=object)
        records = []
        ref = None
        for i, xx in enumerate(x):
            for j, yy in enumerate(y):
                p = provider(float(xx), float(yy))
                if not isinstance(p, SampledPSF):
                    raise TypeError("A PSF provider must return SampledPSF")
                ref = ref or p
                same = (p.intensity.shape == ref.intensity.shape
                        and np.allclose(p.pixel_size_um, ref.pixel_size_um, rtol=1e-10,
                                        atol=1e-12)
                        and np.isclose(p.wavelength_nm, ref.wavelength_nm, rtol=1e-10,
                                       atol=1e-12))
                if not same:
                    raise ValueError("Provider kernels must share shape, pitch and wavelength; "
                                     "resample explicitly")
                grid[i, j] = p.intensity
                records.append(dict(x_um=float(xx), y_um=float(yy), provider=p.provenance))
        kernels = np.stack([np.stack(list(row)) for row in grid])
        return cls(kernels, x, y, ref.pixel_size_um, ref.wavelength_nm,
                   {"samples": records,
                    "interpolation": "bilinear at source; nearest edge extension"})

    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, intensity=self.intensity, x_um=self.x_um, y_um=self.y_um,
                            pixel_size_um=self.pixel_size_um, wavelength_nm=self.wavelength_nm,
                            axes=_FIELD_AXES, normalization="unit_sum",
                            interpolation=_FIELD_INTERP, provenance=json.dumps(self.provenance))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as d:
            if (str(d["axes"]) != _FIELD_AXES or str(d["normalization"]) != "unit_sum"
                    or str(d["interpolation"]) != _FIELD_INTERP):
                raise ValueError("Unsupported PSF field axes, normalization or interpolation")
            return cls(d["intensity"], d["x_um"], d["y_um"], tuple(d["pixel_size_um"]),
                       float(d["wavelength_nm"]), json.loads(str(d["provenance"])))


def _device(device):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("maomi_amd.psf needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device(device)


def resample_psf(psf, pixel_size_um, shape, *, device="cuda"):
    """Explicit linear intensity resampling about the kernel origin (maomi/psf.py:154-176):
    zero outside the supplied support, then renormalised to unit sum; the mass before the
    renormalisation is recorded in the provenance.  An approximation, not optical
    propagation; throughput is specified separately."""
    import torch
    from . import _native as nat
    pitch = np.asarray(pixel_size_um, dtype=float)
    if pitch.shape != (2,) or not np.isfinite(pitch).all() or (pitch <= 0).any():
        raise ValueError("Specify two positive specimen-plane pixel pitches")
    if len(shape) != 2 or any(int(n) != n or n < 2 for n in shape):
        raise ValueError("shape must contain two integer dimensions >= 2")
    shape = tuple(int(n) for n in shape)
    # sample coordinates in source pixels, per axis, in the reference's operation order
    axes = [(np.arange(n) - n // 2) * step / old + size // 2
            for n, step, old, size in zip(shape, pitch, psf.pixel_size_um, psf.intensity.shape)]
    dev = _device(device)
    src = torch.as_tensor(np.ascontiguousarray(psf.intensity), device=dev)
    cx, cy = (torch.as_tensor(np.ascontiguousarray(c, np.float64), device=dev) for c in axes)
    out = torch.empty(shape, dtype=torch.float64, device=dev)
    sx, sy = psf.intensity.shape
    nat.check(nat.lib().maomi_resample_linear_f64(nat.ptr(src), sx, sy, nat.ptr(cx), shape[0],
                                                  nat.ptr(cy), shape[1], nat.ptr(out),
                                                  nat.stream_handle(dev)), "resample psf")
    a = out.cpu().numpy()
    mass = float(a.sum() * np.prod(pitch) / np.prod(psf.pixel_size_um))
    return SampledPSF.normalized(
        a, tuple(pitch), psf.wavelength_nm,
        {"source": psf.provenance,
         "resampling": "linear intensity; zero extension; unit-sum renormalization",
         "source_pixel_size_um": psf.pixel_size_um, "source_shape": list(psf.intensity.shape),
         "pre_normalization_mass_estimate": mass})


def pupil_to_intensity(opd_waves, fft_size=256, amplitude=None, *, device="cuda"):
    """Scalar diffraction intensity of a pupil given its OPD in waves (maomi/psf.py:179-203).

    NaN marks blocked or untraced rays; a finite zero OPD is a valid ray.  The pupil, odd or
    even, is embedded with its discrete origin at fft_size // 2; the result is |FFT|^2 in the
    centred layout, normalised to unit sum (FP64 on the GPU)."""
    import torch
    from . import _native as nat
    opd = np.asarray(opd_waves, dtype=float)
    if opd.ndim != 2 or min(opd.shape) < 2:
        raise ValueError("OPD must be a 2-D grid")
    if int(fft_size) != fft_size or fft_size < max(opd.shape):
        raise ValueError("fft_size must be an integer no smaller than the pupil grid")
    amp = np.ones_like(opd) if amplitude is None else np.asarray(amplitude, dtype=float)
    if amp.shape != opd.shape or not np.isfinite(amp).all() or (amp < 0).any():
        raise ValueError("Pupil amplitude must match the OPD and be finite and nonnegative")
    if not (np.isfinite(opd) & (amp > 0)).any():
        raise ValueError("No transmitting pupil samples")
    n = int(fft_size)
    dev = _device(device)
    st = nat.stream_handle(dev)
    lib = nat.lib()
    o = torch.as_tensor(np.ascontiguousarray(opd), device=dev)
    a = torch.as_tensor(np.ascontiguousarray(amp), device=dev)
    field_ = torch.empty((n, n), dtype=torch.complex128, device=dev)
    nat.check(lib.maomi_pupil_embed_f64(nat.ptr(o), nat.ptr(a), opd.shape[0], opd.shape[1], n,
                                        nat.ptr(field_), st), "pupil embed")
    spec = torch.fft.fft2(field_).contiguous()
    inten = torch.empty((n, n), dtype=torch.float64, device=dev)
    nat.check(lib.maomi_power_shift_f64(nat.ptr(spec), n, nat.ptr(inten), st), "power shift")
    inten /= inten.sum()
    return inten.cpu().numpy()
