"""Optical configuration, pupil, raw fields and the GPU `propagate` (call surface of maomi/optics.py).

`propagate` keeps the reference's signature and result dictionary
(maomi/optics.py:200-381) but runs every site on the GPU through
OpticalPlan (one persistent launch) and assembles the masks with the FP64
GPU v4.  There is no CPU propagation path in this package.
"""

from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

from . import _optical_math as legacy
from . import _native as nat
from .excitation import excitation_order


@dataclass
class OpticalConfig:
    """Optical defaults and NAOMi parameter derivation (optics.py:21-72)."""
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

    def parameters(self):
        excitation_order(self)
        if not np.isfinite(self.temporal_length_um) or self.temporal_length_um <= 0:
            raise ValueError("Positive temporal length required")
        zw = (np.asarray(self.zernike_weights, dtype=float).ravel()
              if len(self.zernike_weights) else None)
        p = legacy.PsfParams(NA=self.NA, objNA=self.objNA, n=self.n, lambda_exc=self.wavelength_um,
                             obj_fl=self.focal_length_mm, ss=self.ss, psf_sz=np.array(self.psf_um),
                             prop_sz=self.prop_um, taillength=self.tail_um, n_diff=self.n_diff,
                             propcrop=self.propcrop, fastmask=False, zernikeWt=zw)
        v = legacy.VolParams(vol_sz=np.array(self.volume_um), vres=self.vres,
                             vol_depth=self.depth_um, verbose=0)
        v.vasc_sz = (legacy.gaussian_beam_size(p, self.depth_um + self.volume_um[2] / 2)
                     + v.vol_sz + np.array([0, 0, self.depth_um]))
        return v, p


def locations(volume_um, count_per_axis):
    """count x count sample points over the FOV incl. edges; one sample = FOV centre."""
    axes = [np.linspace(0, volume_um[i], count_per_axis) if count_per_axis > 1
            else np.array([volume_um[i] / 2]) for i in (0, 1)]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 2)


def _pupil_grid(config):
    v, p = config.parameters()
    fl = float(config.focal_length_mm / 1000.0)
    D2 = float(1e-6 * (1.0 / config.vres) / config.ss)
    N = np.float32(1e-6 * (v.vasc_sz[:2] - v.vol_sz[:2]) / D2)
    D1 = float(max(legacy.gaussian_beam_size(p, fl * 1e6)) / 1e6 / min(N))
    return v, p, N, D1, fl


def pupil(config, *, custom_phase=None, phase_units="radians", phase_aperture="objective",
          device="cpu"):
    """Pupil field with explicit phase units and aperture (optics.py:90-117).

    On a ROCm device the Gaussian/Zernike pupil comes from the HIP kernel
    maomi_pupil_gaussian; on "cpu" from the torch generate_ba mirror.
    """
    v, p, N, D1, fl = _pupil_grid(config)
    if custom_phase is not None:
        p.zernikeWt = None
    dev = torch.device(device)
    if dev.type == "cuda" and p.type == "gaussian" and int(N[0]) == int(N[1]) and N[0] == int(N[0]):
        n = int(N[0])
        u = torch.empty((n, n), dtype=torch.complex64, device=dev)
        zw = generate = legacy.generate_zernike(p)
        abb = np.zeros(0) if isinstance(generate, int) else np.asarray(zw, float)
        nz = int(len(abb)) if np.any(abb != 0) else 0
        zt = torch.as_tensor(abb.astype(np.float32), device=dev) if nz else None
        zmap = torch.empty((n, n), dtype=torch.float32, device=dev) if nz else None
        dwork = torch.zeros(1, dtype=torch.float64, device=dev)
        rad = float(np.tan(np.arcsin(p.NA / p.n)) * fl)
        objrad = float(np.tan(np.arcsin(p.objNA / p.n)) * fl)
        k = 2 * float(p.n) * np.pi / float(p.lambda_exc * 1e-6)
        nat.check(nat.lib().maomi_pupil_gaussian(nat.ptr(u), n, D1, rad, objrad, k, fl, nat.ptr(zt),
                                                 nz, nat.ptr(zmap), nat.ptr(dwork),
                                                 nat.stream_handle(dev)), "pupil")
    else:
        u = legacy.generate_ba(v, p, device=dev)
    if custom_phase is None:
        return u
    if phase_units not in ("radians", "waves") or phase_aperture not in ("objective", "excitation"):
        raise ValueError("Explicit phase_units / phase_aperture required")
    h, w = u.shape
    d1 = max(legacy.gaussian_beam_size(p, fl * 1e6)) / 1e6 / min(h, w)
    na = config.objNA if phase_aperture == "objective" else config.NA
    rad = np.tan(np.arcsin(na / config.n)) * fl
    yy, xx = torch.meshgrid(torch.arange(-h / 2, h / 2, device=dev) * d1,
                            torch.arange(-w / 2, w / 2, device=dev) * d1, indexing="ij")
    grid = torch.stack((xx / rad, yy / rad), -1)[None]
    ph = torch.as_tensor(custom_phase, dtype=u.real.dtype, device=dev)[None, None]
    ph = F.grid_sample(ph, grid, mode="nearest", align_corners=True)[0, 0]
    return u * torch.exp(1j * ph * (2 * np.pi if phase_units == "waves" else 1))


def _shapes(config):
    v, p = config.parameters()
    r = config.vres
    n = int(round((v.vasc_sz[0] - v.vol_sz[0]) * r))
    n2 = (int(round(min(max(legacy.gaussian_beam_size(p, config.psf_um[2] / 2, 3)) * r * 2, n)))
          if config.propcrop else n)
    za = int(round((config.depth_um + config.volume_um[2] / 2 - config.psf_um[2] / 2) * r))
    zb = int(round(config.psf_um[2] * r))
    pc = int(round(config.prop_um * r))
    nc = int(round(config.tail_um * r))
    xy = np.round(np.array(config.volume_um[:2]) * r).astype(int)
    return n, n2, za, zb, pc, nc, xy


def prepare_fields(config, vessels=None, *, device="cpu", scatter=True, out_device=None,
                   events=None):
    """Seeded raw A/B/C index fields (optics.py:120-197); input generation, untimed.

    The noise is drawn with torch's generator on `device`, as the reference does; synthesis
    runs on the GPU (maomi_amd.inputs).  Fields are returned on `out_device` (default
    `device`)."""
    from .inputs import prepare_fields as gpu_fields
    return gpu_fields(config, vessels, device=device, scatter=scatter, out_device=out_device,
                      shapes=_shapes(config), events=events)


def propagate(config, raw_fields, points_um, *, uin=None, device="cuda", save_full=False,
              phase_mode="auto", phase_memory_limit_mb=256):
    """Absolute PSFs, raw/relative masks and tails for explicit sample points (optics.py:200-381).

    Runs on the GPU: all points in one persistent launch, FP64 v4 assembly.
    save_full and the phase mode are accepted for signature compatibility; the
    GPU path always saves only the output crops (numerically identical).
    """
    if not str(device).startswith("cuda"):
        raise ValueError("maomi_amd.propagate runs on a ROCm GPU; pass device='cuda'")
    if phase_mode not in ("auto", "full", "window"):
        raise ValueError("phase_mode must be auto/full/window")
    from .batched import OpticalPlan
    from .fixed import assemble
    u = pupil(config, device=device) if uin is None else torch.as_tensor(uin, device=device)
    if u.dtype != torch.complex64:
        u = u.to(torch.complex64)
    plan = OpticalPlan(config, raw_fields, uin=u, device=device)
    pts = np.asarray(points_um, float)
    if np.any(pts < 0) or np.any(pts > np.array(config.volume_um[:2])):
        raise ValueError("PSF coordinates outside FOV")
    data = plan.evaluate(pts, tails=True, tail_convmask="mean")
    shape = tuple(np.round(np.array(config.volume_um[:2]) * config.vres).astype(int))
    res = assemble(config, data, shape, mask_mode="matlab_v4", z_positions=plan.z_positions,
                   cut=plan.cut)
    estimate = sum(int(np.prod(a.shape)) for a in raw_fields) * config.ss ** 2 * 8
    res.update(mask_mode="explicit_full_paths_v4", scattering_replay="provided_raw_fields",
               phase_mode="full", estimated_full_phase_bytes=int(estimate))
    return res


def collection_mask(vessels, config, events=None):
    """Depth-dependent collection disk and hemoglobin absorption (optics.py:384-415); input
    generation, on the GPU (maomi_amd.inputs.collection_mask)."""
    from .inputs import collection_mask as gpu_mask
    return gpu_mask(vessels, config, events=events)


def run_optics(config, points_um, *, vessels=None, scatter=True, custom_phase=None,
               phase_units="radians", phase_aperture="objective", device="cuda", rana=False):
    """prepare_fields + pupil + propagate + collection mask (optics.py:418-458)."""
    fields = prepare_fields(config, vessels, scatter=scatter, device="cpu")
    u = pupil(config, custom_phase=custom_phase, phase_units=phase_units,
              phase_aperture=phase_aperture, device=device)
    out = propagate(config, fields, points_um, uin=u, device=device)
    out["colmask"] = (collection_mask(vessels, config) if vessels is not None
                      else np.ones_like(out["mask"]))
    out["settings"] = {"seed": config.seed, "scatter": scatter, "phase_units": phase_units,
                       "phase_aperture": phase_aperture, "low_order_subtraction": False}
    if rana is not False and rana is not None:
        raise NotImplementedError("RANA acquisition is outside this package's optical stage")
    return out
