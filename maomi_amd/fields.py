"""Sampled local-PSF fields and the PSF-provider boundary.

Same contracts as the reference's RANA types, so that fields built here load there and back:
  * FullFieldPSF      maomi/rana/fields.py:17-110  (validation, from_provider, NPZ schema
                      "maomi.full_field_psf.v1");
  * DepthPSFBank      maomi/rana/depth.py:9-69     (NPZ schema "maomi.depth_psf_bank.v1");
  * PSFProvider, resolve_field  maomi/rana/providers.py:10-40  (any object with build_field();
                      resolved once at preparation; errors propagate, no shared-PSF fallback).
These are host-side containers (NumPy, FP64).
"""

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Protocol, runtime_checkable

import numpy as np

_FULL_SCHEMA = "maomi.full_field_psf.v1"
_BANK_SCHEMA = "maomi.depth_psf_bank.v1"


def _nodes(c, what):
    a = np.array(c, dtype=float, copy=True)
    if a.ndim != 1 or len(a) < 2 or not np.isfinite(a).all() or (np.diff(a) <= 0).any():
        raise ValueError(what)
    return a


@dataclass(frozen=True)
class FullFieldPSF:
    """Nonnegative effective-intensity kernels on an XY grid of field nodes.

    kernels: axes field_X, field_Y, kernel_X, kernel_Y[, kernel_Z]; kernel origin at
    shape // 2; x_um, y_um: increasing node coordinates (at least two per axis); one voxel
    pitch per kernel axis (micrometres); operator 'detection' or 'excitation'.  Arrays are
    copied to FP64 and made read-only; provenance must be JSON without NaN."""

    kernels: np.ndarray
    x_um: np.ndarray
    y_um: np.ndarray
    voxel_size_um: tuple
    wavelength_nm: float
    operator: str
    provenance: dict = field(default_factory=dict)

    def __post_init__(self):
        if np.iscomplexobj(self.kernels):
            raise ValueError("Supply effective real intensity, not a complex optical field")
        k = np.array(self.kernels, dtype=np.float64, copy=True)
        msg = "Full-field PSFs require at least two increasing finite nodes on EACH XY axis"
        x, y = _nodes(self.x_um, msg), _nodes(self.y_um, msg)
        if (k.ndim not in (4, 5) or k.shape[:2] != (len(x), len(y)) or min(k.shape[2:]) < 1
                or not np.isfinite(k).all() or (k < 0).any()):
            raise ValueError("Supply nonnegative finite kernels with axes "
                             "field_X,field_Y,kernel_X,kernel_Y[,kernel_Z]")
        energy = k.sum(axis=tuple(range(2, k.ndim)))
        if not np.isfinite(energy).all() or (energy <= 0).any():
            raise ValueError("Every local PSF must have positive finite energy")
        pitch = tuple(float(v) for v in self.voxel_size_um)
        if len(pitch) != k.ndim - 2 or not np.isfinite(pitch).all() or min(pitch) <= 0:
            raise ValueError("Specify one positive voxel pitch per kernel dimension")
        if not np.isfinite(self.wavelength_nm) or self.wavelength_nm <= 0:
            raise ValueError("wavelength_nm must be positive and finite")
        if self.operator not in ("detection", "excitation"):
            raise ValueError("Declare operator='detection' or operator='excitation' explicitly")
        meta = json.loads(json.dumps(self.provenance, allow_nan=False))
        for a in (k, x, y):
            a.setflags(write=False)
        for name, value in (("kernels", k), ("x_um", x), ("y_um", y), ("voxel_size_um", pitch),
                            ("provenance", meta)):
            object.__setattr__(self, name, value)

    @classmethod
    def from_provider(cls, provider, x_um, y_um, *, voxel_size_um, wavelength_nm, operator,
                      provenance=None):
        """Sample provider(x_um, y_um) -> effective kernel on the node grid, unchanged."""
        msg = "Provider grid needs at least two increasing finite nodes per axis"
        x, y = _nodes(x_um, msg), _nodes(y_um, msg)
        k = np.array([[np.asarray(provider(float(a), float(b))) for b in y] for a in x])
        return cls(k, x, y, voxel_size_um, wavelength_nm, operator, provenance or {})

    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, schema=_FULL_SCHEMA, kernels=self.kernels, x_um=self.x_um,
                            y_um=self.y_um, voxel_size_um=self.voxel_size_um,
                            wavelength_nm=self.wavelength_nm, operator=self.operator,
                            provenance=json.dumps(self.provenance, allow_nan=False))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as d:
            if str(d["schema"]) != _FULL_SCHEMA:
                raise ValueError("Unsupported full-field PSF schema")
            return cls(d["kernels"], d["x_um"], d["y_um"], tuple(d["voxel_size_um"]),
                       float(d["wavelength_nm"]), str(d["operator"]),
                       json.loads(str(d["provenance"])))


@dataclass(frozen=True)
class DepthPSFBank:
    """One complete XYZ-kernel FullFieldPSF per focal depth (>= 2 increasing depths)."""

    depths_um: np.ndarray
    fields: tuple

    def __post_init__(self):
        z = np.array(self.depths_um, dtype=float, copy=True)
        fs = tuple(self.fields)
        if (z.ndim != 1 or len(z) < 2 or not np.isfinite(z).all() or (np.diff(z) <= 0).any()
                or (z < 0).any()):
            raise ValueError("A depth bank requires >= 2 increasing nonnegative finite depths")
        if len(fs) != len(z) or any(not isinstance(f, FullFieldPSF) or f.kernels.ndim != 5
                                    for f in fs):
            raise ValueError("One complete XYZ-kernel FullFieldPSF per depth is required")
        f0 = fs[0]
        for f in fs[1:]:
            same = (f.kernels.shape == f0.kernels.shape and f.voxel_size_um == f0.voxel_size_um
                    and f.operator == f0.operator and f.wavelength_nm == f0.wavelength_nm
                    and np.array_equal(f.x_um, f0.x_um) and np.array_equal(f.y_um, f0.y_um))
            if not same:
                raise ValueError("Depth fields must share shape, XY nodes, pitch, operator and "
                                 "wavelength")
        z.setflags(write=False)
        object.__setattr__(self, "depths_um", z)
        object.__setattr__(self, "fields", fs)

    @property
    def voxel_size_um(self):
        return self.fields[0].voxel_size_um

    @property
    def operator(self):
        return self.fields[0].operator

    @property
    def wavelength_nm(self):
        return self.fields[0].wavelength_nm

    @classmethod
    def from_provider(cls, provider, x_um, y_um, depths_um, *, voxel_size_um, wavelength_nm,
                      operator="excitation", provenance=None):
        """provider(x_um, y_um, depth_um) -> kernel; one FullFieldPSF per depth."""
        fields = []
        for z in depths_um:
            fields.append(FullFieldPSF.from_provider(
                lambda a, b, z=float(z): provider(a, b, z), x_um, y_um,
                voxel_size_um=voxel_size_um, wavelength_nm=wavelength_nm, operator=operator,
                provenance={**(provenance or {}), "focal_depth_um": float(z)}))
        return cls(depths_um, tuple(fields))

    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        f0 = self.fields[0]
        np.savez_compressed(path, schema=_BANK_SCHEMA, depths_um=self.depths_um,
                            kernels=np.stack([f.kernels for f in self.fields]), x_um=f0.x_um,
                            y_um=f0.y_um, voxel_size_um=f0.voxel_size_um,
                            wavelength_nm=f0.wavelength_nm, operator=f0.operator,
                            provenance=json.dumps([f.provenance for f in self.fields],
                                                  allow_nan=False))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as d:
            if str(d["schema"]) != _BANK_SCHEMA:
                raise ValueError("Unsupported depth-bank schema")
            meta = json.loads(str(d["provenance"]))
            n = len(d["depths_um"])
            if len(meta) != n or len(d["kernels"]) != n:
                raise ValueError("Malformed depth bank: inconsistent depth, kernel or provenance "
                                 "counts")
            fields = tuple(FullFieldPSF(k, d["x_um"], d["y_um"], tuple(d["voxel_size_um"]),
                                        float(d["wavelength_nm"]), str(d["operator"]), m)
                           for k, m in zip(d["kernels"], meta))
            return cls(d["depths_um"], fields)


@runtime_checkable
class PSFProvider(Protocol):
    """Anything with build_field() -> FullFieldPSF | DepthPSFBank; no registration."""

    def build_field(self): ...


def resolve_field(source):
    """A field passes through; a provider is built once here (never in a frame loop)."""
    if isinstance(source, (FullFieldPSF, DepthPSFBank)):
        return source
    build = getattr(source, "build_field", None)
    if not callable(build):
        raise TypeError("Supply FullFieldPSF, DepthPSFBank or a PSFProvider with build_field()")
    out = build()
    if not isinstance(out, (FullFieldPSF, DepthPSFBank)):
        raise TypeError("PSFProvider.build_field() must return FullFieldPSF or DepthPSFBank")
    return out
