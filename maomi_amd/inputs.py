"""GPU input generation for the optical stage.

* `grf_from_noise` / `masked_gp`: the spectral Gaussian-process field of
  maomi/_optical_math.py:20-67 -- a sum of Gaussian spectral envelopes times complex white
  noise, inverse-FFT'd, real part scaled by 2 p_scale sqrt(N).
* `prepare_fields`: the seeded raw A/B/C index fields with partial-slab thickness and vessel
  occupancy (maomi/optics.py:120-197).
* `collection_mask`: the depth-dependent collection disk and hemoglobin absorption
  (maomi/optics.py:384-415).

Normal draws come from torch's generator on the requested device, in the reference's order
(real part, then imaginary part, field by field, after `manual_seed(config.seed)`), so a seed
gives the reference's draws on that device.  The envelope, the inverse FFT (rocFFT) and the
finishing pass run on the GPU; the collection disk is an exact FP64 run-sum convolution on the
GPU instead of a float32 FFT convolution.  Vessels are read slab by slab through indexing only,
so lazy (on-disk) arrays work.
"""

import numpy as np
import torch

from . import _native as nat


def _gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("maomi_amd input generation needs a ROCm GPU; none is visible")
    nat.lib()
    return torch.device("cuda", torch.cuda.current_device())


def _envelope_tables(dims, l_scale, l_weights, device):
    """1-D envelope factors exp(-linspace(-pi/2, pi/2)^2 * l^2) per scale and axis, and
    (prod l, w^2) per scale, as the reference's float32 tensor expressions
    (_optical_math.py:23-52)."""
    ls = torch.as_tensor(l_scale, dtype=torch.float32, device=device)
    if ls.ndim != 2 or ls.shape[1] != 3 or ls.shape[0] < 1:
        raise ValueError("l_scale must have shape (K, 3)")
    w = (torch.ones(ls.shape[0], device=device) if l_weights is None
         else torch.as_tensor(l_weights, dtype=torch.float32, device=device))
    sq = [torch.linspace(-np.pi / 2, np.pi / 2, d, device=device, dtype=torch.float32).pow(2)
          for d in dims]
    rows = [torch.cat([torch.exp(-sq[a] * ls[i, a] ** 2) for a in range(3)])
            for i in range(ls.shape[0])]
    coef = [torch.stack([torch.prod(ls[i]), w[i] ** 2]) for i in range(ls.shape[0])]
    return torch.stack(rows), torch.stack(coef)


def _spectrum(re, im, l_scale, l_weights, table_device):
    dev = _gpu()
    dims = tuple(int(s) for s in re.shape)
    if len(dims) != 3 or tuple(im.shape) != dims:
        raise ValueError("noise_re and noise_im must be equal 3-D arrays")
    ker, coef = _envelope_tables(dims, l_scale, l_weights, table_device)
    ker = ker.to(dev).contiguous()
    coef = coef.to(dev).contiguous()
    re = torch.as_tensor(re, dtype=torch.float32).to(dev).contiguous()
    im = torch.as_tensor(im, dtype=torch.float32).to(dev).contiguous()
    spec = torch.empty(dims, dtype=torch.complex64, device=dev)
    nat.check(nat.lib().maomi_grf_spectrum(nat.ptr(re), nat.ptr(im), *dims, nat.ptr(ker),
                                           nat.ptr(coef), int(ker.shape[0]), nat.ptr(spec),
                                           nat.stream_handle()), "maomi_grf_spectrum")
    return torch.fft.ifftn(spec, dim=(0, 1, 2)).contiguous()


def _finish(g, dims, scale, mu=0.0, first=1.0, last=1.0, occupancy=None, occ_div=1.0):
    dev = _gpu()
    out = torch.empty(dims, dtype=torch.float32, device=dev)
    occ = None if occupancy is None else occupancy.to(dev, torch.float32).contiguous()
    nat.check(nat.lib().maomi_grf_finish(nat.ptr(g), *dims, float(scale), float(mu),
                                         float(first), float(last), nat.ptr(occ),
                                         float(occ_div), nat.ptr(out), nat.stream_handle()),
              "maomi_grf_finish")
    return out


def grf_from_noise(noise_re, noise_im, l_scale, p_scale, mu=0, bin_mask=1, l_weights=None,
                   table_device="cpu"):
    """Spectral Gaussian-process field from given normal draws (_optical_math.py:20-67).

    Returns a float32 tensor on the GPU.  `table_device` is where the 1-D envelope factors are
    evaluated (the device the reference would evaluate them on)."""
    dims = tuple(int(s) for s in noise_re.shape)
    g = _spectrum(noise_re, noise_im, l_scale, l_weights, table_device)
    N = int(np.prod(dims))
    if np.ndim(bin_mask) == 0 and not torch.is_tensor(bin_mask):
        scale = np.float32(2 * p_scale * bin_mask * np.sqrt(N))
        return _finish(g, dims, scale, mu)
    # an array-valued mask multiplies elementwise before the offset, as in the reference
    real = _finish(g, dims, 1.0)
    mask = torch.as_tensor(bin_mask, device=real.device)
    return 2 * p_scale * mask * float(np.sqrt(N)) * real + mu


def masked_gp(grid_sz, l_scale, p_scale, mu=0, bin_mask=1, l_weights=None, device="cpu"):
    """`masked_3DGP_test` (_optical_math.py:20-67): draws the complex noise with torch's
    generator on `device`, synthesises on the GPU and returns the field on `device`."""
    dims = [int(x) for x in grid_sz]
    re = torch.randn(dims, device=device, dtype=torch.float32)
    im = torch.randn(dims, device=device, dtype=torch.float32)
    return grf_from_noise(re, im, l_scale, p_scale, mu, bin_mask, l_weights,
                          table_device=device).to(device)


def _reader(vessels, dev):
    def read(xs, ys, zs):
        a = vessels[xs, ys, zs]
        if isinstance(a, torch.Tensor):
            return a.to(device=dev, dtype=torch.float32).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(torch.float32).contiguous()
    return read


def _zsum_into(slab, out, zo):
    X, Y, L = (int(s) for s in slab.shape)
    nat.check(nat.lib().maomi_zsum(nat.ptr(slab), X, Y, L, nat.ptr(out), int(out.shape[2]),
                                   int(zo), nat.stream_handle()), "maomi_zsum")


def _occupancy(vessels, shapes, n, n2, za, zb, pc, nc, dev):
    """Vessel occupancy of the A/B/C fields (optics.py:167-190): slab sums for A and C, the
    centred crop for B."""
    read = _reader(vessels, dev)
    need = (shapes[0][0], shapes[0][1], za + zb + nc)
    X, Y = slice(0, need[0]), slice(0, need[1])
    occA = torch.empty(shapes[0], dtype=torch.float32, device=dev)
    bounds = [0] + list(range(za % pc or pc, za, pc)) + [za]
    for k, (s, e) in enumerate(zip(bounds[:-1], bounds[1:])):
        _zsum_into(read(X, Y, slice(s, e)), occA, k)
    off = (n - n2) // 2
    occB = read(slice(off, off + shapes[1][0]), slice(off, off + shapes[1][1]),
                slice(za, za + zb))
    occC = torch.empty(shapes[2], dtype=torch.float32, device=dev)
    for k, z in enumerate(range(za + zb, za + zb + nc, pc)):
        _zsum_into(read(X, Y, slice(z, min(z + pc, za + zb + nc))), occC, k)
    return occA, occB, occC


def prepare_fields(config, vessels=None, *, device="cpu", scatter=True, out_device=None,
                   shapes=None):
    """Seeded raw A/B/C index fields (optics.py:120-197) synthesised on the GPU.

    `device` is where the noise is drawn (the reference's `device`); the fields are returned on
    `out_device` (default: `device`).  `shapes` is (n, n2, za, zb, pc, nc, xy) of
    `maomi_amd.optics._shapes`."""
    dev = _gpu()
    n, n2, za, zb, pc, nc, xy = shapes
    fshapes = [(int(xy[0] + n), int(xy[1] + n), int(np.ceil(za / pc))),
               (int(xy[0] + n2), int(xy[1] + n2), int(zb)),
               (int(xy[0] + n), int(xy[1] + n), int(np.ceil(nc / pc)))]
    occ = [None, None, None]
    if vessels is not None:
        if not hasattr(vessels, "shape"):
            vessels = np.asarray(vessels)
        need = (fshapes[0][0], fshapes[0][1], za + zb + nc)
        if any(a < b for a, b in zip(vessels.shape, need)):
            raise ValueError(f"Padded vessels need at least {need}, got {tuple(vessels.shape)}")
        occ = list(_occupancy(vessels, fshapes, n, n2, za, zb, pc, nc, dev))
    first = [np.float32((za % pc) / pc) if za % pc else 1.0, 1.0, 1.0]
    last = [1.0, 1.0, np.float32((nc % pc) / pc) if nc % pc else 1.0]
    divs = [pc, 1, pc]
    spectra = [None, None, None]
    if scatter:
        rng_devs = [torch.device(device).index or 0] if str(device).startswith("cuda") else []
        with torch.random.fork_rng(devices=rng_devs):
            torch.manual_seed(config.seed)
            r = config.vres
            scales = torch.tensor(config.scatter_scales_um, device=device,
                                  dtype=torch.float32)[:, None].repeat(1, 3) * r
            wts = torch.tensor(config.scatter_weights, device=device, dtype=torch.float32)
            for i, shape in enumerate(fshapes):
                ls = scales.clone()
                ls[:, 2] /= 1 if i == 1 else pc
                re = torch.randn(shape, device=device, dtype=torch.float32)
                im = torch.randn(shape, device=device, dtype=torch.float32)
                spectra[i] = (_spectrum(re, im, ls, wts, device), ls)
    out = []
    for i, shape in enumerate(fshapes):
        N = int(np.prod(shape))
        p_scale = 1 if i == 1 else pc
        g = spectra[i][0] if spectra[i] is not None else None
        scale = np.float32(2 * p_scale * 1 * np.sqrt(N))
        out.append(_finish(g, shape, scale, 0.0, first[i], last[i], occ[i], divs[i]))
    target = device if out_device is None else out_device
    return [a.to(target) for a in out]


def _disk_runs(X, Y, radius):
    """The reference's full-grid disk x^2 + y^2 <= radius^2 on arange(-X/2, X/2) by
    arange(-Y/2, Y/2), cropped to its bounding box (optics.py:396-405); returns the first and
    last column of each row, the cropped size and the cell count, or None when empty."""
    x = np.arange(-X / 2, X / 2)[:, None]
    y = np.arange(-Y / 2, Y / 2)[None, :]
    disk = x * x + y * y <= radius * radius
    rows, cols = np.flatnonzero(disk.any(1)), np.flatnonzero(disk.any(0))
    if not len(rows):
        return None
    d = disk[rows[0]:rows[-1] + 1, cols[0]:cols[-1] + 1]
    K0, K1 = d.shape
    first = d.argmax(1)
    last = K1 - 1 - d[:, ::-1].argmax(1)
    count = d.sum(1)
    if np.any(count != last - first + 1) or np.any(count == 0):
        raise AssertionError("collection disk rows must be single runs")
    runs = np.stack([first, last], 1).astype(np.int32).ravel()
    return runs, K0, K1, int(count.sum())


def collection_mask(vessels, config):
    """Depth-dependent collection disk and hemoglobin absorption (optics.py:384-415).

    For every slab of prop_um, the vessel projection is convolved with a normalised disk whose
    radius follows the collection cone, 'same' placement as scipy.signal.fftconvolve; the sum
    over slabs, cropped to the volume, gives 10 ** (-total / vres * hemoabs).  The convolution
    is an exact FP64 sum of row runs over prefix sums on the GPU.  Returns float32 NumPy."""
    dev = _gpu()
    a = vessels if hasattr(vessels, "shape") else np.asarray(vessels)
    X, Y, Zv = (int(s) for s in (a.shape[0], a.shape[1], a.shape[-1]))
    r = config.vres
    pc = int(round(config.prop_um * r))
    shape = np.round(np.array(config.volume_um[:2]) * r).astype(int)
    ox, oy = ((np.array((X, Y)) - shape) // 2).tolist()
    Xo, Yo = int(shape[0]), int(shape[1])
    if ox < 0 or oy < 0:
        raise ValueError(f"Vessels of {X} x {Y} do not cover the {Xo} x {Yo} volume")
    read = _reader(a, dev)
    total = torch.zeros((Xo, Yo), dtype=torch.float64, device=dev)
    pre = torch.empty((X, Y + 1), dtype=torch.float64, device=dev)
    lib, stream = nat.lib(), nat.stream_handle()
    cone = np.tan(np.arcsin(config.objNA / config.n))
    for start in range(0, Zv, pc):
        depth = (start / pc + 0.5) * config.prop_um
        radius = r * cone * (config.depth_um + config.volume_um[2] / 2 - depth)
        if radius <= 0:
            continue
        disk = _disk_runs(X, Y, radius)
        if disk is None:
            continue
        runs, K0, K1, count = disk
        weight = float(np.float32(1.0) / np.float32(count))
        slab = read(slice(None), slice(None), slice(start, start + pc))
        runs_d = torch.from_numpy(runs).to(dev)
        nat.check(lib.maomi_slab_prefix(nat.ptr(slab), X, Y, int(slab.shape[2]), nat.ptr(pre),
                                        stream), "maomi_slab_prefix")
        nat.check(lib.maomi_disk_accumulate(nat.ptr(pre), X, Y, nat.ptr(runs_d), K0,
                                            (K0 - 1) // 2, (K1 - 1) // 2, weight, ox, oy, Xo,
                                            Yo, nat.ptr(total), stream),
                  "maomi_disk_accumulate")
        del slab, runs_d
    out = torch.empty((Xo, Yo), dtype=torch.float32, device=dev)
    nat.check(lib.maomi_hemo_mask(nat.ptr(total), Xo * Yo, float(np.float32(r)),
                                  float(np.float32(config.hemoabs)), nat.ptr(out), stream),
              "maomi_hemo_mask")
    return out.cpu().numpy()
