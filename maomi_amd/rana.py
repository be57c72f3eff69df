"""Acquisition through a field of local PSFs on the GPU, and the adapter that attaches it to an
optical-stage result.

``RanaOperator(field, *, native=False, origin_um=(0, 0), z_origin_um=0, device='cuda')``
keeps the reference's operator (maomi/rana/operator.py:20-170): smooth XY node weights,
linear focal-depth weights of a DepthPSFBank, ``static(volume)`` and
``scan_slab(slab, z_start=0, scan_avg=1)``, a bounded cache of LocalFFTPlans, nonnegative
results.  The projections run on ``maomi_amd._local_fft.LocalFFTPlan`` (HIP kernels plus
rocFFT, FP64).  ``attach_rana(optics, rana=True, ...)`` follows
maomi/rana/adapters/naomi.py:9-68: with no field it builds one from the run's own per-site
PSFs, otherwise it resolves the given field or provider once.
"""

from collections import OrderedDict
from dataclasses import dataclass, replace
from collections.abc import Mapping

import numpy as np
import torch

from .fields import DepthPSFBank, FullFieldPSF, resolve_field
from .providers import native_field


def axis_weights(nodes, positions):
    """Each position weighs its two bracketing nodes by a smoothstep of its fraction between
    them (operator.py:10-17); the weights of a position sum to one."""
    nodes = np.asarray(nodes, float)
    positions = np.asarray(positions, float)
    k = np.clip(np.searchsorted(nodes, positions, side="right") - 1, 0, len(nodes) - 2)
    f = np.clip((positions - nodes[k]) / (nodes[k + 1] - nodes[k]), 0.0, 1.0)
    f = f * f * (3.0 - 2.0 * f)
    w = np.zeros((len(nodes), len(positions)))
    cols = np.arange(len(positions))
    w[k, cols] = 1.0 - f
    w[k + 1, cols] = f
    return w


class RanaOperator:
    def __init__(self, field, *, native=False, origin_um=(0.0, 0.0), z_origin_um=0.0,
                 device="cuda"):
        field = resolve_field(field)
        self.bank = field if isinstance(field, DepthPSFBank) else None
        self.fields = field.fields if self.bank is not None else (field,)
        if any(not isinstance(f, FullFieldPSF) or f.kernels.ndim != 5 for f in self.fields):
            raise TypeError("The acquisition add-on requires complete XYZ-kernel fields")
        if str(device) == "cpu":
            raise ValueError("this RanaOperator runs on the GPU: pass device='cuda[:index]'")
        self.field = self.fields[0]
        self.kernel_shape = self.field.kernels.shape[2:]
        self.voxel_size_um = self.field.voxel_size_um
        self.origin_um = np.asarray(origin_um, float)
        self.z_origin_um = float(z_origin_um)
        self.native, self.device = bool(native), str(device)
        if (self.origin_um.shape != (2,) or not np.isfinite(self.origin_um).all()
                or not np.isfinite(self.z_origin_um)):
            raise ValueError("Supply finite XY and Z origins")
        self._dev = torch.device(self.device)
        self._weights_cache = {}
        self._plans = OrderedDict()
        self._scan_kernels = {}
        self.settings = {
            "enabled": True, "operator": "native_scan" if native else self.field.operator,
            "interpolation": "positive local cubic tensor weights",
            "field_nodes": list(self.field.kernels.shape[:2]),
            "kernel_shape": list(self.kernel_shape),
            "depth_nodes_um": None if self.bank is None else self.bank.depths_um.tolist(),
            "source_masks": "existing excitation and collection masks retained",
            "native_gain": "source excitation mask" if native else "custom kernel gain plus source masks",
            "tail_resampling": "nonnegative intensity projection, integrated gain preserved",
            "shared_psf_fallback": False, "device": self.device,
            "provenance": [f.provenance for f in self.fields]}

    # ------------------------------------------------------------ weights
    def _weights(self, shape):
        shape = tuple(shape[:2])
        if shape not in self._weights_cache:
            pitch = np.array(self.voxel_size_um[:2])
            lo = np.array([self.field.x_um[0], self.field.y_um[0]])
            hi = np.array([self.field.x_um[-1], self.field.y_um[-1]])
            if ((self.origin_um < lo - 1e-10).any()
                    or (self.origin_um + np.array(shape) * pitch > hi + 1e-10).any()):
                raise ValueError("RANA field must cover the entire acquisition FOV")
            centres = [o + (np.arange(n) + 0.5) * p for o, n, p in zip(self.origin_um, shape, pitch)]
            self._weights_cache[shape] = tuple(
                axis_weights(nodes, c) for nodes, c in zip((self.field.x_um, self.field.y_um),
                                                           centres))
        return self._weights_cache[shape]

    def _depth_weights(self, z):
        if self.bank is None:
            return np.ones((1, np.size(z)))
        z = np.atleast_1d(z)
        d = self.bank.depths_um
        if (z < d[0] - 1e-10).any() or (z > d[-1] + 1e-10).any():
            raise ValueError("RANA depth bank does not cover the requested acquisition depths")
        return np.stack([np.interp(z, d, row) for row in np.eye(len(self.fields))])

    # ------------------------------------------------------------ plans
    def clear_cache(self):
        """Release prepared kernel FFTs and device-side interpolation weights."""
        self._plans.clear()
        self._scan_kernels.clear()
        self._weights_cache.clear()

    @property
    def cache_info(self):
        names = ("kernel_fft_builds", "cache_hits")
        return dict(plans=len(self._plans),
                    fft_bytes=sum(p.cache_bytes for p in self._plans.values()),
                    **{k: sum(p.stats[k] for p in self._plans.values()) for k in names})

    def _project(self, source, kernels, *, plane=False):
        from ._local_fft import LocalFFTPlan
        key = (tuple(source.shape), id(kernels), plane)
        if key not in self._plans:
            limit = max(2, 2 * len(self.fields))
            while len(self._plans) >= limit:
                self._plans.popitem(last=False)[1].clear()
            op = self.field.operator
            self._plans[key] = LocalFFTPlan(
                source.shape, kernels, self._weights(source.shape), plane=plane,
                detection=not self.native and op == "detection",
                correlate=not self.native and op == "excitation", device=self.device,
                cache_mb=256 / limit)
        self._plans.move_to_end(key)
        return self._plans[key](source)

    def _array(self, a):
        return torch.as_tensor(np.asarray(a, np.float64), device=self._dev)

    @staticmethod
    def _finish(result):
        return np.maximum(result.cpu().numpy(), 0.0)

    @staticmethod
    def physical_tail(values):
        """Clip negative ringing of a tail intensity, keeping its integrated gain."""
        a = np.asarray(values)
        if np.iscomplexobj(a) or not np.isfinite(a).all() or a.sum(dtype=float) <= 0:
            raise ValueError("Tail intensity must have positive finite integrated gain")
        if np.any(a < 0):
            pos = np.maximum(a, 0)
            return pos * (a.sum(dtype=float) / pos.sum(dtype=float))
        return a

    @staticmethod
    def _density(volume):
        if np.iscomplexobj(volume):
            raise ValueError("Fluorescence must be real")
        a = np.asarray(volume, dtype=float)
        if a.ndim != 3 or not np.isfinite(a).all() or (a < 0).any():
            raise ValueError("Supply finite nonnegative XYZ fluorescence")
        return a

    # ------------------------------------------------------------ acquisition
    def static(self, volume):
        """Static XYZ image of a fluorescence volume through the local kernels."""
        a = self._density(volume)
        z = self.z_origin_um + (np.arange(a.shape[2]) + 0.5) * self.voxel_size_um[2]
        depth = self._depth_weights(z)
        src = self._array(a)
        out = torch.zeros_like(src)
        detection = not self.native and self.field.operator == "detection"
        for f, w in zip(self.fields, depth):
            if not np.any(w):
                continue
            wd = self._array(w)
            if detection:
                out += self._project(src * wd, f.kernels)
            else:
                out += self._project(src, f.kernels) * wd
        return self._finish(out)

    def scan_slab(self, slab, *, z_start=0, scan_avg=1):
        """One scanned plane of a focal slab (depth equal to the kernel depth)."""
        a = self._density(slab)
        if a.shape[2] != self.kernel_shape[2]:
            raise ValueError("Focal slab depth must match the RANA kernel")
        if not isinstance(scan_avg, (int, np.integer)) or scan_avg < 1:
            raise ValueError("scan_avg must be a positive integer")
        Z = a.shape[2]
        centre = self.kernel_shape[2] // 2
        z = self.z_origin_um + (z_start + np.arange(Z) + 0.5) * self.voxel_size_um[2]
        detection = not self.native and self.field.operator == "detection"
        depth = self._depth_weights(z if detection else [z[centre]])
        starts = np.arange(0, Z, scan_avg) if scan_avg > 1 else None
        src = self._array(a)
        out = torch.zeros(a.shape[:2], dtype=torch.float64, device=self._dev)
        for index, (f, w) in enumerate(zip(self.fields, depth)):
            if not np.any(w):
                continue
            s = src * self._array(w) if detection else src
            key = (index, scan_avg, detection)
            if key not in self._scan_kernels:
                k = f.kernels
                if detection:  # kernel depth reversed about its centre, zero past the end
                    ids = 2 * centre - np.arange(Z)
                    k = np.take(k, np.minimum(ids, Z - 1), axis=-1).copy()
                    k[..., ids >= Z] = 0
                if starts is not None:
                    k = np.add.reduceat(k, starts, axis=-1)
                self._scan_kernels[key] = k
            if starts is not None:
                s = torch.stack([s[..., b:b + scan_avg].sum(-1) for b in starts], dim=-1)
            proj = self._project(s.contiguous(), self._scan_kernels[key], plane=True)
            out += proj if detection else proj * float(w[0])
        return self._finish(out)


@dataclass(frozen=True)
class RanaAddon:
    """field=None keeps the run's own per-position PSFs; otherwise a FullFieldPSF,
    DepthPSFBank or provider, or a per-channel mapping of them (rana/config.py:69-83)."""

    field: object = None
    voxel_size_um: tuple = None
    origin_um: tuple = (0.0, 0.0)
    z_origin_um: float = 0.0
    device: str = None
    status_window: bool = True  # no effect here: there is no viewer window


def attach_rana(optics, rana=True, *, config=None, volume=None, device=None):
    """The optics mapping plus an acquisition operator (and 'image' when a volume is given);
    rana False/None returns the mapping unchanged."""
    if rana is False or rana is None:
        return optics
    addon = RanaAddon() if rana is True else rana if isinstance(rana, RanaAddon) else RanaAddon(field=rana)
    dev = addon.device or device or "cuda"
    if "channels" in optics:
        if isinstance(addon.field, Mapping) and set(addon.field) != set(optics["channels"]):
            raise ValueError("Supply one RANA field per optical channel")
        chans = {name: attach_rana(
            value, replace(addon, field=addon.field[name] if isinstance(addon.field, Mapping)
                           else addon.field), config=config, volume=volume, device=dev)
            for name, value in optics["channels"].items()}
        out = dict(optics, channels=chans)
        if volume is not None:
            out["image"] = sum(optics["beam_weights"][n] * v["image"] for n, v in chans.items())
        out["settings"] = dict(optics.get("settings", {}), rana={"enabled": True,
                                                               "channels": list(chans)})
        return out
    native = addon.field is None
    pitch = addon.voxel_size_um
    if pitch is None and config is not None:
        pitch = (1 / config.vres,) * 3
    if native:
        if pitch is None:
            raise ValueError("Specify voxel_size_um when attaching native PSFs without OpticalConfig")
        field = native_field(optics, config=config, voxel_size_um=pitch,
                             wavelength_nm=config.wavelength_um * 1000 if config is not None else 920.0)
    else:
        field = resolve_field(addon.field)
    if pitch is not None and not np.allclose(field.voxel_size_um, pitch, rtol=1e-10, atol=1e-12):
        raise ValueError("RANA PSF voxel size must match the acquisition volume")
    op = RanaOperator(field, native=native, origin_um=addon.origin_um,
                      z_origin_um=addon.z_origin_um, device=dev)
    out = dict(optics, rana=op)
    out["settings"] = dict(optics.get("settings", {}), rana=op.settings)
    if volume is not None:
        a = np.asarray(volume, float)
        for key in ("mask", "colmask"):
            if optics.get(key) is not None:
                a = a * np.asarray(optics[key])[..., None]
        out["image"] = op.static(a)
    return out
