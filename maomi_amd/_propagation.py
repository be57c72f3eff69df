"""Propagation geometry shared by the stage drivers (call surface of maomi/_propagation.py).

geometry() restates the reference's integer and float32 rules exactly: field
sides n and n2, pupil and sample pitches d1 and d2, plane positions of the
A/B/C segments (partial first slab when za is not a multiple of the slab),
the first top-tail plane `cut`, and the centred crops.
"""

from types import SimpleNamespace

import numpy as np
import torch

from . import _optical_math as math


def center_slice(n, m):
    """Centred m-window of an n-grid (_propagation.py:9-12)."""
    if m > n or (n - m) % 2:
        raise ValueError(f"Incompatible centered grids {n},{m}")
    return slice((n - m) // 2, (n + m) // 2)


def geometry(config, raw_fields):
    """_propagation.py:15-56.  raw_fields: the three raw A/B/C arrays (only depths are read)."""
    v, p = config.parameters()
    r, ss = config.vres, config.ss
    n = int(round((v.vasc_sz[0] - v.vol_sz[0]) * r * ss))
    if config.propcrop:
        n2 = int(round(min(max(math.gaussian_beam_size(p, config.psf_um[2] / 2, 3)) * r * 2 * ss, n)))
    else:
        n2 = n
    if n % 2 and not n2 % 2:
        n2 -= 1
    fl = np.float32(config.focal_length_mm / 1000)
    d1 = np.float32(max(math.gaussian_beam_size(p, float(fl * 1e6))) / 1e6 / n)
    d2 = np.float32(1e-6 / r / ss)
    za = r * (config.depth_um + config.volume_um[2] / 2 - config.psf_um[2] / 2)
    pc = config.prop_um * r
    a, b, c = [f.shape[-1] for f in raw_fields]
    frac = za % pc
    z_a = (np.r_[0, np.arange(a) + frac / pc] if frac else np.arange(a + 1)) * config.prop_um * 1e-6
    z_b = np.arange(b + 1) / r * 1e-6
    z_c = np.r_[np.arange(c) * config.prop_um, config.tail_um] * 1e-6
    cut = max(0, int(np.ceil((za - config.tail_um * r) / pc)))
    if cut >= len(z_a) - 1:
        raise ValueError("Top tail has no extent")
    focal = np.round(np.array(config.psf_um[:2]) * r * ss).astype(int)
    return SimpleNamespace(
        n=n, n2=n2, fl=fl, d1=d1, d2=d2,
        wvl=np.float32(config.wavelength_um * 1e-6), nr=np.float32(config.n),
        z=[z_a, z_b, z_c], cut=cut, focal_shape=focal,
        crop=center_slice(n, n2), focal_crop=tuple(center_slice(n2, int(k)) for k in focal),
    )


def edge_apodization(size, device, dtype):
    """exp(-a^16) x exp(-a^16), a = (i - size/2)/(0.47 size) (_propagation.py:59-61)."""
    a = torch.arange(-size / 2, size / 2, device=device, dtype=dtype) / (0.47 * size)
    return torch.exp(-a[:, None] ** 16) * torch.exp(-a[None, :] ** 16)
