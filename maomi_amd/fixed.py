"""Fixed-lattice optical stage: sites, field shapes, mask assembly, run_fixed (call surface of maomi/fixed.py).

assemble() builds the MATLAB-v4 excitation masks from per-site energies and
tail weights with the FP64 GPU v4 (Green build, dense solve, Green apply) and
keeps the reference's result dictionary (maomi/fixed.py:119-190).  run_fixed()
is the timed pipeline of maomi/fixed.py:193-411 on one ROCm GPU: plan (pupil +
OpticalPlan), propagation (one persistent launch for all sites with tails),
aggregation (assemble); static scanning stays outside the timer.
"""

import time

import numpy as np
import torch

from . import _native as nat

from .optics import collection_mask, prepare_fields, pupil, _shapes


def fixed_locations(config, count_per_axis=65):
    """count x count physical locations snapped to the supersampled lattice (fixed.py:15-33)."""
    if isinstance(count_per_axis, bool) or not isinstance(count_per_axis, int) or count_per_axis < 2:
        raise ValueError("count_per_axis must be an integer >= 2")
    ticks = np.asarray(config.volume_um[:2], float) * config.vres * config.ss
    if (not np.all(np.isfinite(ticks)) or np.any(ticks < count_per_axis - 1)
            or not np.allclose(ticks, np.round(ticks))):
        raise ValueError(
            f"Fixed {count_per_axis}x{count_per_axis} requires an integer optical lattice "
            f"with at least {count_per_axis} positions per axis; choose count_per_axis explicitly")
    axes = [np.round(np.linspace(0, int(round(t)), count_per_axis)) / (config.vres * config.ss)
            for t in ticks]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 2)
    if len(np.unique(pts, axis=0)) != count_per_axis ** 2:
        raise ValueError("Grid contains duplicate locations")
    return pts


def field_shapes(config):
    """Raw A/B/C field shapes (fixed.py:36-56)."""
    n, n2, za, zb, pc, nc, xy = _shapes(config)
    if min(za, zb, pc, nc) <= 0:
        raise ValueError("Optical A/B/C propagation extents must be positive")
    return [(xy[0] + n, xy[1] + n, int(np.ceil(za / pc))), (xy[0] + n2, xy[1] + n2, zb),
            (xy[0] + n, xy[1] + n, int(np.ceil(nc / pc)))]


def assemble(config, out, shape, *, mask_mode="matlab_v4", z_positions=None, cut=None,
             device=None, v4_ahead=None):
    """Masks and PSF statistics from native outputs (fixed.py:119-190); v4 on the GPU in FP64.
    v4_ahead: a numerics.V4Ahead started on these points (times vres), whose point-only part of
    the solve is then taken instead of computed here."""
    from .numerics import v4_gpu
    if mask_mode == "cubic":
        raise ValueError("mask_mode 'cubic' is a legacy diagnostic outside this package")
    if mask_mode != "matlab_v4":
        raise ValueError("mask_mode must be cubic or matlab_v4")
    points, psfs = out["points_um"], out["psfs"]
    stats = out.get("psf_stats")  # device-side sums from OpticalPlan.evaluate, when present
    energies = stats["energies"] if stats is not None else psfs.sum((1, 2, 3), dtype=np.float64)
    tails = bool(out["tails_computed"])
    values = [energies] + ([out[k]["weight"] for k in ("top", "bottom")] if tails else [])
    values = np.stack(values, axis=1)
    q = np.stack(np.meshgrid(*[np.arange(k) + 0.5 for k in shape], indexing="ij"), -1)
    if len(points) == 1:
        masks = np.broadcast_to(values[0], tuple(shape) + values.shape[1:]).copy()
    else:
        dev = torch.device(device or "cuda")
        prep = v4_ahead.take() if v4_ahead is not None else None
        masks = v4_gpu(points * config.vres, values, q.reshape(-1, 2), dev,
                       prepared=prep).cpu().numpy()
        masks = masks.reshape(tuple(shape) + (values.shape[1],))
    weights = np.full(len(points), 1 / len(points))
    mask = masks[..., 0]
    mean = float(mask.mean())
    if not np.isfinite(mean) or mean <= 0:
        raise ValueError("Nonpositive or invalid mean excitation energy")
    if stats is not None:
        P = len(points)
        psf_mean = (stats["sum_abs"] / P).reshape(psfs.shape[1:])
        psf_raw = (stats["sum"] / P).reshape(psfs.shape[1:])
        neg = float(-stats["sum_neg"].sum() / max(stats["sum_abs"].sum(), 1e-30))
    else:
        flat = psfs.reshape(len(points), -1)
        psf_mean = (weights @ np.abs(flat)).reshape(psfs.shape[1:])
        psf_raw = (weights @ flat).reshape(psfs.shape[1:])
        neg = float(-np.minimum(psfs, 0).sum(dtype=np.float64)
                    / max(np.abs(psfs).sum(dtype=np.float64), 1e-30))
    result = {
        "psf": psf_mean,
        "psf_raw": psf_raw,
        "psf_3d": psfs, "points_um": points, "energies": energies,
        "mask_raw": mask, "mask": mask / mean, "mask_mean": mean, "mask_mode": mask_mode,
        "geometry_cache_hit": False, "tails_computed": tails,
        "unique_optical_paths": len(points),
        "negative_resampling_mass_fraction": neg,
    }
    if tails:
        for j, (native, name) in enumerate([("top", "psfT"), ("bottom", "psfB")]):
            data = out[native]
            profile = data["profile"].sum(0, dtype=np.float64)
            z = z_positions[0][cut:] if j == 0 else z_positions[2]
            z = z - z[0]
            axial = np.arange(round(z[-1] * 1e6 * config.vres) + 1) / config.vres
            interp = np.interp(axial, z * 1e6, profile)
            tm = masks[..., j + 1]
            tmean = float(tm.mean())
            result[name] = {
                "convmask": (data["convmask_mean"] if "convmask_mean" in data
                             else data["convmask"].mean(0)),
                "weight": float(data["weight"].mean()),
                "psfZ": interp / interp.mean() if interp.mean() > 0 else interp,
                "mask_raw": tm, "mask": tm / tmean if tmean > 0 else tm, "mask_mean": tmean,
                "weights_per_point": data["weight"], "profiles_per_point": data["profile"],
            }
    return result


def zstack(volume, psf, *, mask=None, collection_mask=None, device="cuda"):
    """MATLAB convn(volume * mask * collection_mask, psf, 'same') with zero extension and the
    MATLAB crop floor(k/2) on every axis (scanning.py:8-53, mode matlab_convn), in FP64.

    One tile covering the volume through the local-FFT kernels: maomi_lfft_gather applies the
    masks while writing the zero-padded volume, maomi_lfft_product multiplies the spectra and
    maomi_lfft_scatter writes the 'same' window; the rFFTs are rocFFT's.  Returns a float32
    tensor on `device`, as the reference does for float32 input."""
    from scipy.fft import next_fast_len
    dev = torch.device(device)
    v = torch.as_tensor(np.asarray(volume, np.float64), device=dev).contiguous()
    p = np.asarray(psf, np.float64)
    if v.ndim != 3 or p.ndim != 3:
        raise ValueError("Expected X,Y,Z arrays")
    X, Y, Z = v.shape
    w = np.ones((X, Y))
    for m in (mask, collection_mask):
        if m is not None:
            m = np.asarray(m, np.float64)
            if m.shape != (X, Y):
                raise ValueError("Mask must match volume XY")
            w = w * m
    n = tuple(next_fast_len(a + b - 1) for a, b in zip(v.shape, p.shape))
    lib, st = nat.lib(), nat.stream_handle(dev)
    wd = torch.as_tensor(w.ravel(), device=dev)
    gd = torch.as_tensor(np.array([[0, 0, X, Y, 0]], np.int32), device=dev)
    tiles = torch.empty((1,) + n, dtype=torch.float64, device=dev)
    nat.check(lib.maomi_lfft_gather(nat.ptr(v), X, Y, Z, nat.ptr(gd), nat.ptr(wd), nat.ptr(tiles),
                                    1, n[0], n[1], n[2], st), "zstack gather")
    spec = torch.fft.rfftn(tiles, dim=(1, 2, 3)).contiguous()
    kspec = torch.fft.rfftn(torch.as_tensor(p, device=dev)[None], s=n, dim=(1, 2, 3)).contiguous()
    nat.check(lib.maomi_lfft_product(nat.ptr(spec), nat.ptr(kspec), None, spec.numel(), 0, st),
              "zstack product")
    full = torch.fft.irfftn(spec, s=n, dim=(1, 2, 3)).contiguous()
    out = torch.zeros((X, Y, Z), dtype=torch.float64, device=dev)
    c = [b // 2 for b in p.shape]
    sd = torch.as_tensor(np.array([[0, 0, X, Y, c[0], c[1], c[2], -1]], np.int32), device=dev)
    nat.check(lib.maomi_lfft_scatter(nat.ptr(full), 1, n[0], n[1], n[2], nat.ptr(sd), None,
                                     nat.ptr(out), X, Y, Z, X, Y, st), "zstack crop")
    return out.float()


def run_fixed(config, volume, *, vessels=None, raw_fields=None, scatter=True, custom_phase=None,
              phase_units="radians", phase_aperture="objective", input_pupil=None,
              beam_weights=None, tails=True, batch_size=32, count_per_axis=65,
              mask_mode="matlab_v4", device="cuda", rana=False, render_image=True):
    """MATLAB v4 masks and arithmetic PSF averaging on a count x count lattice (fixed.py:193-411)."""
    from .batched import OpticalPlan
    if rana is not False and rana is not None:
        # native run without an image, then the acquisition operator (fixed.py:219-232); the
        # reference's status window has no counterpart here
        from .rana import attach_rana
        overall = time.perf_counter()
        native = run_fixed(config, volume, vessels=vessels, raw_fields=raw_fields,
                           scatter=scatter, custom_phase=custom_phase, phase_units=phase_units,
                           phase_aperture=phase_aperture, input_pupil=input_pupil,
                           beam_weights=beam_weights, tails=tails, batch_size=batch_size,
                           count_per_axis=count_per_axis, mask_mode=mask_mode, device=device,
                           rana=False, render_image=False)
        start = time.perf_counter()
        result = attach_rana(native, rana, config=config,
                             volume=volume if render_image else None, device=device)
        result["timing_seconds"] = dict(native.get("timing_seconds", {}),
                                        rana=time.perf_counter() - start)
        result["native_total_seconds"] = native["total_seconds"]
        result["total_seconds"] = time.perf_counter() - overall
        return result
    if mask_mode != "matlab_v4":
        raise ValueError("run_fixed requires matlab_v4; use maomi.enhanced.CubicFullFieldRenderer "
                         "with a complete custom FullFieldPSF for cubic imaging")
    points = fixed_locations(config, count_per_axis)
    if isinstance(input_pupil, dict):
        if custom_phase is not None:
            raise ValueError("Compose channel pupil phases explicitly")
        if not input_pupil:
            raise ValueError("At least one beam required")
        weights = {k: 1.0 for k in input_pupil} if beam_weights is None else dict(beam_weights)
        if (set(weights) != set(input_pupil)
                or not all(np.isfinite(w) and w >= 0 for w in weights.values())
                or sum(weights.values()) <= 0):
            raise ValueError("Nonnegative detected-intensity weight per beam required")
        start = time.perf_counter()
        raw = (prepare_fields(config, vessels, scatter=scatter, out_device=device)
               if raw_fields is None else raw_fields)
        channels = {name: run_fixed(config, volume, vessels=vessels, raw_fields=raw,
                                    input_pupil=f, tails=tails, batch_size=batch_size,
                                    count_per_axis=count_per_axis, mask_mode=mask_mode,
                                    device=device, render_image=render_image)
                    for name, f in input_pupil.items()}
        return {
            "channels": channels, "beam_weights": weights,
            **({"image": sum(weights[k] * v["image"] for k, v in channels.items())}
               if render_image else {}),
            "points_um": points, "unique_optical_paths": len(points) * len(channels),
            "unique_optical_positions": len(points), "total_seconds": time.perf_counter() - start,
            "settings": {"combination": "incoherent detected-intensity sum",
                         "scaling": config.scaling, "paths_per_position": len(channels),
                         "count_per_axis": count_per_axis, "mask_mode": mask_mode},
        }
    if beam_weights is not None:
        raise ValueError("beam_weights requires a mapping of pupils")
    if not str(device).startswith("cuda"):
        raise ValueError("Fixed fast pipeline requires a GPU; the CPU reference path is separate")
    if not torch.cuda.is_available():
        raise RuntimeError("ROCm PyTorch and an MI355X GPU are required for run_fixed")
    if not isinstance(batch_size, int) or batch_size < 1:
        raise ValueError("batch_size must be a positive integer")
    if phase_units not in ("radians", "waves") or phase_aperture not in ("objective", "excitation"):
        raise ValueError("Invalid pupil phase units/aperture")
    if input_pupil is not None and custom_phase is not None:
        raise ValueError("Supply either input_pupil or custom_phase; compose the field explicitly")
    expected = tuple(np.round(np.asarray(config.volume_um) * config.vres).astype(int))
    if render_image:
        volume = np.asarray(volume, np.float32)
        if volume.shape != expected or not np.isfinite(volume).all():
            raise ValueError(f"Expected finite volume of shape {expected}")
    started = time.perf_counter()
    timings = {}
    supplied = raw_fields is not None
    t = time.perf_counter()
    raw = (prepare_fields(config, vessels, scatter=scatter, out_device=device)
           if raw_fields is None else raw_fields)
    exp_shapes = field_shapes(config)
    if len(raw) != 3 or any(tuple(a.shape) != tuple(s) for a, s in zip(raw, exp_shapes)):
        raise ValueError(f"Raw A/B/C fields must have shapes {exp_shapes}")
    if any(torch.is_complex(torch.as_tensor(a)) or not bool(torch.isfinite(torch.as_tensor(a)).all())
           for a in raw):
        raise ValueError("Raw refractive-index fields must be finite and real")
    col = collection_mask(vessels, config) if vessels is not None else np.ones(expected[:2], np.float32)
    timings["fields_and_collection"] = time.perf_counter() - t
    with torch.cuda.device(device):
        t = time.perf_counter()
        # the point-only part of the v4 solve (Green matrix, projection, Cholesky) depends on the
        # lattice alone: queued first on a side stream, it runs beside the plan stage
        ahead = None
        if len(points) > 1:
            from .numerics import V4Ahead
            ahead = V4Ahead(torch.device(device))
            ahead.start(np.asarray(points, np.float64) * config.vres)
        u = (pupil(config, custom_phase=custom_phase, phase_units=phase_units,
                   phase_aperture=phase_aperture, device=device)
             if input_pupil is None
             else torch.as_tensor(input_pupil, dtype=torch.complex64, device=device))
        n_exp = int(round((config.parameters()[0].vasc_sz[0] - config.volume_um[0])
                          * config.vres * config.ss))
        if (tuple(u.shape) != (n_exp, n_exp) or not bool(torch.isfinite(u).all())
                or not bool(torch.any(u.abs() > 0))):
            raise ValueError("Input pupil must be a finite nonzero complex field on the optical grid")
        plan = OpticalPlan(config, raw, uin=u, device=device)
        torch.cuda.synchronize()
        timings["plan"] = time.perf_counter() - t
        t = time.perf_counter()
        # the PSF stack's copy to the host runs beside the v4 solve; resolved after assemble
        outputs = plan.evaluate(points, batch_size=batch_size, tails=tails, tail_convmask="mean",
                                defer_psfs=True)
        timings["propagation"] = time.perf_counter() - t
        if plan.last_launch_seconds() is not None:
            timings["propagation_kernel"] = plan.last_launch_seconds()
        t = time.perf_counter()
        result = assemble(config, outputs, expected[:2], mask_mode=mask_mode,
                          z_positions=plan.z_positions, cut=plan.cut, device=device,
                          v4_ahead=ahead)
        if hasattr(result["psf_3d"], "result"):
            result["psf_3d"] = result["psf_3d"].result()
        result["colmask"] = col
        timings["aggregation"] = time.perf_counter() - t
        t = time.perf_counter()
        if render_image:
            result["image"] = zstack(volume, result["psf"].astype(np.float32),
                                     mask=result["mask"].astype(np.float32), collection_mask=col,
                                     device=device).cpu().numpy()
        timings["static_scan"] = time.perf_counter() - t
    result["timing_seconds"] = timings
    result["total_seconds"] = time.perf_counter() - started
    result["settings"] = {
        "seed": config.seed, "scaling": config.scaling,
        "fields": "provided" if supplied else "generated",
        "scatter": None if supplied else bool(scatter), "phase_units": phase_units,
        "phase_aperture": phase_aperture,
        "input_pupil": "provided" if input_pupil is not None else "gaussian",
        "collection_absorption": vessels is not None, "tail_background_added_to_image": False,
        "count_per_axis": count_per_axis, "adaptive": False, "low_order_subtraction": False,
        "mask_mode": mask_mode, "psf_average": "arithmetic mean of absolute per-point PSFs",
    }
    return result
