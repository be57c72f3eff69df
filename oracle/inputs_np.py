"""CPU restatement of MaoMi's input generation -- TEST INFRASTRUCTURE ONLY.

* masked_gp: maomi/_optical_math.py:20-67 (spectral Gaussian-process field) given the two
  normal draws, FP64 envelope and inverse FFT.
* prepare_fields: maomi/optics.py:120-197 (raw A/B/C fields, partial first/last slab
  thickness, vessel occupancy) given the draws of each field.
* collection_mask: maomi/optics.py:384-415 (collection disk and hemoglobin absorption) with
  scipy.signal.fftconvolve in FP64.

Pinned in tests/test_oracle_inputs.py: a single-frequency draw gives the closed-form plane
wave (shift direction, grid, envelope, 2 p sqrt(N) scale); envelopes of several scales add in
power; the field variance equals 4 p^2 mean(envelope^2); a uniform vessel slab gives the
closed-form absorption inside the volume; a single vessel voxel gives the disk footprint of
weight 1/count; a brute-force direct 'same' convolution equals the FFT one.
"""

import numpy as np
from scipy.signal import fftconvolve

from .optics_np import field_shapes


def envelope(dims, l_scale, l_weights=None):
    """sqrt(sum_i prod(l_i) w_i^2 prod_axes exp(-g_a l_ia^2)^2), g = linspace(-pi/2, pi/2)^2
    (_optical_math.py:23-52), FP64."""
    ls = np.asarray(l_scale, np.float64)
    w = np.ones(len(ls)) if l_weights is None else np.asarray(l_weights, np.float64)
    g = [np.linspace(-np.pi / 2, np.pi / 2, d) ** 2 for d in dims]
    acc = np.zeros(dims)
    for i in range(len(ls)):
        k = (np.exp(-g[0] * ls[i, 0] ** 2)[:, None, None]
             * np.exp(-g[1] * ls[i, 1] ** 2)[None, :, None]
             * np.exp(-g[2] * ls[i, 2] ** 2)[None, None, :])
        acc += np.prod(ls[i]) * k ** 2 * w[i] ** 2
    return np.sqrt(acc)


def masked_gp(re, im, l_scale, p_scale, mu=0.0, bin_mask=1.0, l_weights=None):
    """_optical_math.py:20-67: 2 p sqrt(N) bin_mask Re(ifftshift(ifftn(ifftshift(E (re + i im)))))
    + mu."""
    re = np.asarray(re, np.float64)
    dims = re.shape
    spec = envelope(dims, l_scale, l_weights) * (re + 1j * np.asarray(im, np.float64))
    g = np.fft.ifftshift(np.fft.ifftn(np.fft.ifftshift(spec)))
    return 2 * p_scale * bin_mask * np.sqrt(re.size) * g.real + mu


def _shape_parts(config):
    shapes = field_shapes(config)
    r = config.vres
    za = int(round((config.depth_um + config.volume_um[2] / 2 - config.psf_um[2] / 2) * r))
    zb = int(round(config.psf_um[2] * r))
    pc = int(round(config.prop_um * r))
    nc = int(round(config.tail_um * r))
    return shapes, za, zb, pc, nc


def prepare_fields(config, noises, vessels=None, scatter=True):
    """optics.py:120-197.  noises: per field (re, im) draws in the field's shape."""
    shapes, za, zb, pc, nc = _shape_parts(config)
    r = config.vres
    scales = np.repeat(np.asarray(config.scatter_scales_um, np.float64)[:, None], 3, 1) * r
    raw = []
    for i, shape in enumerate(shapes):
        ls = scales.copy()
        slab = 1 if i == 1 else pc
        ls[:, 2] /= slab
        if scatter:
            re, im = noises[i]
            raw.append(masked_gp(re, im, ls, slab, l_weights=config.scatter_weights))
        else:
            raw.append(np.zeros(shape))
    if za % pc:
        raw[0][..., 0] *= (za % pc) / pc
    if nc % pc:
        raw[2][..., -1] *= (nc % pc) / pc
    if vessels is not None:
        v = np.asarray(vessels, np.float64)
        X, Y = shapes[0][0], shapes[0][1]
        bounds = [0] + list(range(za % pc or pc, za, pc)) + [za]
        occA = np.stack([v[:X, :Y, s:e].sum(-1) for s, e in zip(bounds[:-1], bounds[1:])], -1)
        off = (shapes[0][0] - shapes[1][0]) // 2
        occB = v[off:off + shapes[1][0], off:off + shapes[1][1], za:za + zb]
        occC = np.stack([v[:X, :Y, z:min(z + pc, za + zb + nc)].sum(-1)
                         for z in range(za + zb, za + zb + nc, pc)], -1)
        for i, occ in enumerate((occA, occB, occC)):
            raw[i] = occ + raw[i] * (1 - occ / (1 if i == 1 else pc))
    return raw


def collection_mask(vessels, config):
    """optics.py:384-415 in FP64."""
    a = np.asarray(vessels, np.float64)
    r = config.vres
    pc = int(round(config.prop_um * r))
    total = np.zeros(a.shape[:2])
    for start in range(0, a.shape[-1], pc):
        depth = (start / pc + 0.5) * config.prop_um
        radius = r * np.tan(np.arcsin(config.objNA / config.n)) * (
            config.depth_um + config.volume_um[2] / 2 - depth)
        if radius <= 0:
            continue
        x = np.arange(-a.shape[0] / 2, a.shape[0] / 2)[:, None]
        y = np.arange(-a.shape[1] / 2, a.shape[1] / 2)[None, :]
        disk = (x * x + y * y <= radius * radius).astype(np.float64)
        nz = np.argwhere(disk)
        if not len(nz):
            continue
        lo, hi = nz.min(0), nz.max(0) + 1
        disk = disk[lo[0]:hi[0], lo[1]:hi[1]]
        disk /= disk.sum()
        total += fftconvolve(a[..., start:start + pc].sum(-1), disk, mode="same")
    shape = np.round(np.array(config.volume_um[:2]) * r).astype(int)
    o = (np.array(total.shape) - shape) // 2
    total = total[o[0]:o[0] + shape[0], o[1]:o[1] + shape[1]]
    return 10 ** (-total / r * config.hemoabs)
