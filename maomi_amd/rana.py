"""Acquisition through a field of local PSFs on the GPU, and the adapter that attaches it to an
optical-stage result.

``RanaOperator(field, *, native=False, origin_um=(0, 0), z_origin_um=0, device='cuda')`` has
the reference operator's public surface (maomi/rana/operator.py:20-170): ``static(volume)``,
``scan_slab(slab, z_start=0, scan_avg=1)``, ``clear_cache``, ``cache_info``, ``settings``,
``physical_tail``, with the same validation texts and settings keys, and the same results
(smooth node weights in XY, linear weights between the depth fields of a DepthPSFBank,
nonnegative output).  The work is organised for the GPU:
  * every field of a depth bank is one device stack of kernels, uploaded once;
  * node weights (smoothstep between bracketing nodes) and depth weights (hat functions over
    the depth nodes) are computed by HIP kernels (maomi_rana_node_weights,
    maomi_rana_depth_weights) and stay on the device, cached per shape and depth set;
  * one local-FFT plan per (source shape, plane, scan variant) serves every field
    (LocalFFTPlan.apply): the excitation operators transform each source tile once for all
    fields, a plane projection with one weight per field combines the fields' spectra before a
    single inverse transform, and per-z weights are applied inside the gather or scatter
    kernels instead of on weighted copies of the volume;
  * the scan kernels (depth-reversed for detection, summed over scan_avg planes) are built on
    the device from the stack, once per variant.
One upload of the source and one download of the result per call.
``attach_rana(optics, rana=True, ...)`` follows maomi/rana/adapters/naomi.py:9-68: with no field
it builds one from the run's own per-site PSFs, otherwise it resolves the given field or
provider once.
"""

from collections import OrderedDict
from dataclasses import dataclass, replace
from collections.abc import Mapping

import numpy as np
import torch

from . import _native as nat
from ._local_fft import LocalFFTPlan
from .fields import DepthPSFBank, FullFieldPSF, resolve_field
from .providers import native_field

_PLAN_LIMIT = 4  # prepared plans kept (static and scan variants of the shapes in use)


class RanaOperator:
    def __init__(self, field, *, native=False, origin_um=(0.0, 0.0), z_origin_um=0.0,
                 device="cuda"):
        field = resolve_field(field)
        bank = field if isinstance(field, DepthPSFBank) else None
        fields = field.fields if bank is not None else (field,)
        if any(not isinstance(f, FullFieldPSF) or f.kernels.ndim != 5 for f in fields):
            raise TypeError("The acquisition add-on requires complete XYZ-kernel fields")
        if str(device) == "cpu":
            raise ValueError("this RanaOperator runs on the GPU: pass device='cuda[:index]'")
        origin = np.asarray(origin_um, float)
        if origin.shape != (2,) or not np.isfinite(origin).all() or not np.isfinite(z_origin_um):
            raise ValueError("Supply finite XY and Z origins")
        self.bank, self.fields, self.field = bank, fields, fields[0]
        self.kernel_shape = self.field.kernels.shape[2:]
        self.voxel_size_um = self.field.voxel_size_um
        self.origin_um, self.z_origin_um = origin, float(z_origin_um)
        self.native, self.device = bool(native), str(device)
        self._dev = torch.device(self.device)
        op = self.field.operator
        self._detection = not self.native and op == "detection"
        self._correlate = not self.native and op == "excitation"
        # the depth fields as one device stack (F, nx, ny, kx, ky, kz)
        self._stack = torch.as_tensor(np.stack([f.kernels for f in fields]), device=self._dev)
        self._depth_nodes = (None if bank is None else
                             torch.as_tensor(bank.depths_um, dtype=torch.float64,
                                             device=self._dev))
        self._node_w = {}        # XY shape -> host (wx, wy) from the device kernel
        self._depth_w = {}       # z positions -> (F, nz) device weights
        self._scan_stacks = {}   # (scan_avg, detection) -> scan kernel stack
        self._plans = OrderedDict()
        self.settings = {
            "enabled": True, "operator": "native_scan" if native else op,
            "interpolation": "positive local cubic tensor weights",
            "field_nodes": list(self.field.kernels.shape[:2]),
            "kernel_shape": list(self.kernel_shape),
            "depth_nodes_um": None if bank is None else bank.depths_um.tolist(),
            "source_masks": "existing excitation and collection masks retained",
            "native_gain": ("source excitation mask" if native
                            else "custom kernel gain plus source masks"),
            "tail_resampling": "nonnegative intensity projection, integrated gain preserved",
            "shared_psf_fallback": False, "device": self.device,
            "provenance": [f.provenance for f in fields]}

    # ------------------------------------------------------------ weights on the device
    def _xy_weights(self, shape):
        """Node weights of an X x Y acquisition grid (pixel centres from origin_um), computed
        on the device; the plans take a host copy once per shape for their tile layout."""
        shape = tuple(int(n) for n in shape[:2])
        if shape not in self._node_w:
            pitch = np.array(self.voxel_size_um[:2])
            lo = np.array([self.field.x_um[0], self.field.y_um[0]])
            hi = np.array([self.field.x_um[-1], self.field.y_um[-1]])
            if (self.origin_um < lo - 1e-10).any() or \
                    (self.origin_um + np.array(shape) * pitch > hi + 1e-10).any():
                raise ValueError("RANA field must cover the entire acquisition FOV")
            lib, st = nat.lib(), nat.stream_handle(self._dev)
            out = []
            for nodes, o, step, n in zip((self.field.x_um, self.field.y_um), self.origin_um,
                                         pitch, shape):
                nd = torch.as_tensor(nodes, dtype=torch.float64, device=self._dev)
                w = torch.empty((len(nodes), n), dtype=torch.float64, device=self._dev)
                nat.check(lib.maomi_rana_node_weights(nat.ptr(nd), len(nodes), float(o),
                                                      float(step), n, nat.ptr(w), st),
                          "node weights")
                out.append(w)
            self._node_w[shape] = tuple(w.cpu().numpy() for w in out)
        return self._node_w[shape]

    def _z_weights(self, z):
        """(F, nz) device weights of the depth fields at positions z; None for one field."""
        z = np.atleast_1d(np.asarray(z, float))
        if self.bank is None:
            return None
        d = self.bank.depths_um
        if (z < d[0] - 1e-10).any() or (z > d[-1] + 1e-10).any():
            raise ValueError("RANA depth bank does not cover the requested acquisition depths")
        key = z.tobytes()
        if key not in self._depth_w:
            zt = torch.as_tensor(z, device=self._dev)
            w = torch.empty((len(self.fields), len(z)), dtype=torch.float64, device=self._dev)
            nat.check(nat.lib().maomi_rana_depth_weights(nat.ptr(self._depth_nodes),
                                                         len(self.fields), nat.ptr(zt), len(z),
                                                         nat.ptr(w), nat.stream_handle(self._dev)),
                      "depth weights")
            self._depth_w[key] = w
        return self._depth_w[key]

    # ------------------------------------------------------------ plans
    def clear_cache(self):
        """Release prepared kernel FFTs and device-side interpolation weights."""
        self._plans.clear()
        self._scan_stacks.clear()
        self._node_w.clear()
        self._depth_w.clear()

    @property
    def cache_info(self):
        names = ("kernel_fft_builds", "cache_hits")
        return dict(plans=len(self._plans),
                    fft_bytes=sum(p.cache_bytes for p in self._plans.values()),
                    **{k: sum(p.stats[k] for p in self._plans.values()) for k in names})

    def _plan(self, shape, stack, variant, plane):
        key = (tuple(shape), variant, plane)
        if key not in self._plans:
            while len(self._plans) >= _PLAN_LIMIT:
                self._plans.popitem(last=False)[1].clear()
            self._plans[key] = LocalFFTPlan(shape, stack, self._xy_weights(shape), plane=plane,
                                            detection=self._detection,
                                            correlate=self._correlate, device=self.device,
                                            cache_mb=256 / _PLAN_LIMIT)
        self._plans.move_to_end(key)
        return self._plans[key]

    def _scan_stack(self, Z, scan_avg):
        """Kernels of a scan: depth reversed about the centre for detection (zero past the
        end), then summed over groups of scan_avg planes."""
        key = (scan_avg, self._detection)
        if key not in self._scan_stacks:
            k = self._stack
            if self._detection:
                centre = self.kernel_shape[2] // 2
                ids = 2 * centre - np.arange(Z)
                src = torch.as_tensor(np.minimum(ids, Z - 1), device=self._dev)
                k = k.index_select(-1, src)
                k[..., torch.as_tensor(ids >= Z, device=self._dev)] = 0.0
            if scan_avg > 1:
                k = _group_sum(k, scan_avg)
            self._scan_stacks[key] = k.contiguous()
        return self._scan_stacks[key]

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
        """Static XYZ image of a fluorescence volume through the local kernels: every depth
        field in one plan, each weighted by its depth weight along z."""
        a = self._density(volume)
        z = self.z_origin_um + (np.arange(a.shape[2]) + 0.5) * self.voxel_size_um[2]
        zw = self._z_weights(z)
        plan = self._plan(a.shape, self._stack, "static", False)
        return self._finish(plan.apply(torch.as_tensor(a, device=self._dev), zweights=zw))

    def scan_slab(self, slab, *, z_start=0, scan_avg=1):
        """One scanned plane of a focal slab (depth equal to the kernel depth)."""
        a = self._density(slab)
        if a.shape[2] != self.kernel_shape[2]:
            raise ValueError("Focal slab depth must match the RANA kernel")
        if not isinstance(scan_avg, (int, np.integer)) or scan_avg < 1:
            raise ValueError("scan_avg must be a positive integer")
        Z = a.shape[2]
        z = self.z_origin_um + (z_start + np.arange(Z) + 0.5) * self.voxel_size_um[2]
        stack = self._scan_stack(Z, int(scan_avg))
        src = torch.as_tensor(a, device=self._dev)
        if self._detection:
            # per-z depth weights act on the source, before any grouping of planes
            zw = self._z_weights(z)
            if scan_avg > 1:
                F = len(self.fields)
                w = torch.ones((F, Z), dtype=torch.float64, device=self._dev) if zw is None else zw
                srcs = torch.stack([_group_sum(src * w[f], scan_avg) for f in range(F)])
                plan = self._plan(tuple(srcs.shape[1:]), stack, ("scan", scan_avg), True)
                out = sum(plan.apply(srcs[f], fields=[f]) for f in range(F))
                return self._finish(out)
            plan = self._plan(a.shape, stack, ("scan", scan_avg), True)
            return self._finish(plan.apply(src, zweights=zw))
        # excitation / native: one weight per field, at the slab's centre plane
        centre = self.kernel_shape[2] // 2
        zw = self._z_weights([z[centre]])
        s = _group_sum(src, scan_avg) if scan_avg > 1 else src
        plan = self._plan(tuple(s.shape), stack, ("scan", scan_avg), True)
        fw = None if zw is None else zw[:, 0].contiguous()
        return self._finish(plan.apply(s.contiguous(), fweights=fw))


def _group_sum(a, n):
    """Sum of consecutive groups of n planes along the last axis (the last group may be
    shorter), on the device."""
    Z = a.shape[-1]
    groups = torch.as_tensor(np.arange(Z) // n, device=a.device)
    out = torch.zeros(a.shape[:-1] + (-(-Z // n),), dtype=a.dtype, device=a.device)
    return out.index_add_(a.ndim - 1, groups, a)


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
