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
    sq = torch.cat([torch.linspace(-np.pi / 2, np.pi / 2, d, device=device,
                                   dtype=torch.float32).pow(2) for d in dims])
    axis = torch.repeat_interleave(torch.arange(3, device=device),
                                   torch.tensor(dims, device=device))
    # every scale and axis in one expression, element by element the reference's
    # exp(-sq * l^2)
    rows = torch.exp(-sq[None, :] * (ls ** 2)[:, axis])
    coef = [torch.stack([torch.prod(ls[i]), w[i] ** 2]) for i in range(ls.shape[0])]
    return rows, torch.stack(coef)


def _spectrum(re, im, l_scale, l_weights, table_device, events=None):
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
    ev = _record(events, "grf_spectrum", dims)
    nat.check(nat.lib().maomi_grf_spectrum(nat.ptr(re), nat.ptr(im), *dims, nat.ptr(ker),
                                           nat.ptr(coef), int(ker.shape[0]), nat.ptr(spec),
                                           nat.stream_handle()), "maomi_grf_spectrum")
    if ev is not None:
        ev[1].record()
    return torch.fft.ifftn(spec, dim=(0, 1, 2)).contiguous()


def _record(events, name, dims):
    """Start a timing event pair on the current stream when `events` (a list) is given."""
    if events is None:
        return None
    pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    events.append((name, tuple(dims), pair))
    pair[0].record()
    return pair


def _finish(g, dims, scale, mu=0.0, first=1.0, last=1.0, occupancy=None, occ_div=1.0,
            events=None):
    dev = _gpu()
    out = torch.empty(dims, dtype=torch.float32, device=dev)
    occ = None if occupancy is None else occupancy.to(dev, torch.float32).contiguous()
    ev = _record(events, "grf_finish" + ("_occupancy" if occ is not None else ""), dims)
    nat.check(nat.lib().maomi_grf_finish(nat.ptr(g), *dims, float(scale), float(mu),
                                         float(first), float(last), nat.ptr(occ),
                                         float(occ_div), nat.ptr(out), nat.stream_handle()),
              "maomi_grf_finish")
    if ev is not None:
        ev[1].record()
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


def _reader(vessels, dev, raw=False):
    """read(xs, ys, zs) -> contiguous device tensor of that vessel block: float32, or with
    raw=True boolean blocks as their 0/1 bytes (uint8) and anything else as float32."""
    def read(xs, ys, zs):
        a = vessels[xs, ys, zs]
        if not isinstance(a, torch.Tensor):
            a = np.ascontiguousarray(a)
            a = torch.from_numpy(a.view(np.uint8) if a.dtype == np.bool_ else a)
        if raw and a.dtype in (torch.bool, torch.uint8):
            return a.to(dev).contiguous().view(torch.uint8)
        return a.to(device=dev, dtype=torch.float32).contiguous()
    return read


def _elem_bytes(t):
    return 1 if t.dtype == torch.uint8 else 4


def _device_source(vessels):
    """The vessels themselves, as (tensor, element bytes), when they are a contiguous 3-D GPU
    tensor of bool, uint8 or float32 that the kernels can read in place; None otherwise."""
    if (isinstance(vessels, torch.Tensor) and vessels.is_cuda and vessels.dim() == 3
            and vessels.is_contiguous()):
        if vessels.dtype in (torch.bool, torch.uint8):
            return vessels.view(torch.uint8), 1
        if vessels.dtype == torch.float32:
            return vessels, 4
    return None


def _zsum_runs(vessels, X, Y, runs, out, stride_run, stride_xy, dev):
    """Sums of the z runs [(z0, z1), ...] of vessels[:X, :Y] into out (one launch; GPU-resident
    vessels are read in place, others through indexing of the z range the runs cover)."""
    if any(z0 > z1 for z0, z1 in runs) or any(a[1] > b[0] for a, b in zip(runs, runs[1:])):
        raise ValueError("z runs must be sorted and must not overlap")
    src = _device_source(vessels)
    if src is not None:
        block, eb, zoff = src[0][:X, :Y], src[1], 0
    else:
        z_lo, z_hi = min(r[0] for r in runs), max(r[1] for r in runs)
        block = _reader(vessels, dev, raw=True)(slice(0, X), slice(0, Y), slice(z_lo, z_hi))
        eb, zoff = _elem_bytes(block), z_lo
    rd = torch.tensor([[z0 - zoff, z1 - zoff] for z0, z1 in runs], dtype=torch.int32).to(dev)
    nat.check(nat.lib().maomi_zsum(nat.ptr(block), eb, X, Y, int(block.shape[2]),
                                   block.stride(0), block.stride(1), nat.ptr(rd), len(runs),
                                   nat.ptr(out), stride_run, stride_xy, nat.stream_handle()),
              "maomi_zsum")


def _occupancy(vessels, shapes, n, n2, za, zb, pc, nc, dev):
    """Vessel occupancy of the A/B/C fields (optics.py:167-190): slab sums for A and C, the
    centred crop for B."""
    X, Y = shapes[0][0], shapes[0][1]
    occA = torch.empty(shapes[0], dtype=torch.float32, device=dev)
    bounds = [0] + list(range(za % pc or pc, za, pc)) + [za]
    _zsum_runs(vessels, X, Y, list(zip(bounds[:-1], bounds[1:])), occA, 1, shapes[0][2], dev)
    off = (n - n2) // 2
    occB = _reader(vessels, dev)(slice(off, off + shapes[1][0]), slice(off, off + shapes[1][1]),
                                 slice(za, za + zb))
    occC = torch.empty(shapes[2], dtype=torch.float32, device=dev)
    runs = [(z, min(z + pc, za + zb + nc)) for z in range(za + zb, za + zb + nc, pc)]
    _zsum_runs(vessels, X, Y, runs, occC, 1, shapes[2][2], dev)
    return occA, occB, occC


def prepare_fields(config, vessels=None, *, device="cpu", scatter=True, out_device=None,
                   shapes=None, events=None):
    """Seeded raw A/B/C index fields (optics.py:120-197) synthesised on the GPU.

    `device` is where the noise is drawn (the reference's `device`); the fields are returned on
    `out_device` (default: `device`).  `shapes` is (n, n2, za, zb, pc, nc, xy) of
    `maomi_amd.optics._shapes`.  With `events` (a list), a (name, dims, (start, end)) HIP event
    pair is appended for every field-kernel launch."""
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
                spectra[i] = (_spectrum(re, im, ls, wts, device, events), ls)
    out = []
    for i, shape in enumerate(fshapes):
        N = int(np.prod(shape))
        p_scale = 1 if i == 1 else pc
        g = spectra[i][0] if spectra[i] is not None else None
        scale = np.float32(2 * p_scale * 1 * np.sqrt(N))
        out.append(_finish(g, shape, scale, 0.0, first[i], last[i], occ[i], divs[i], events))
    target = device if out_device is None else out_device
    return [a.to(target) for a in out]


def _disk_runs(X, Y, radius):
    """The reference's full-grid disk x^2 + y^2 <= radius^2 on arange(-X/2, X/2) by
    arange(-Y/2, Y/2), cropped to its bounding box (optics.py:396-405); returns the first and
    last column of each row, the cropped size and the cell count, or None when empty."""
    return _disk_runs_many(X, Y, [radius])[0]


def _disk_runs_many(X, Y, radii):
    """_disk_runs for several radii in one vectorised pass.

    Each row of a disk is one run of columns.  Its ends come from the closed form
    y = +-sqrt(radius^2 - x^2), whose rounding moves them by at most one column, and are then
    settled by the reference's own predicate on the columns within two of the estimate, so the
    cells are exactly those of the full grid."""
    x = np.arange(-X / 2, X / 2)
    y0 = -Y / 2
    radii = np.asarray(radii, np.float64)
    sel = np.abs(x)[None, :] <= radii[:, None] + 1            # candidate rows per radius
    slab, row = np.nonzero(sel)
    out = [None] * len(radii)
    if not len(row):
        return out
    xx = x[row] * x[row]
    r2 = radii[slab] * radii[slab]
    h = np.sqrt(np.maximum(r2 - xx, 0.0))
    est_lo = np.clip(np.ceil(-h - y0), 0, Y - 1).astype(np.int64)
    est_hi = np.clip(np.floor(h - y0), 0, Y - 1).astype(np.int64)
    win = np.arange(-2, 3)
    lo_c = est_lo[:, None] + win[None, :]        # ascending: first inside is the run start
    hi_c = est_hi[:, None] - win[None, :]        # descending: first inside is the run end

    def inside(j):
        yj = y0 + j
        return (j >= 0) & (j < Y) & (xx[:, None] + yj * yj <= r2[:, None])

    ok_lo, ok_hi = inside(lo_c), inside(hi_c)
    full = ok_lo.any(1)
    if not np.array_equal(full, ok_hi.any(1)):
        raise AssertionError("collection disk rows must be single runs")
    lo = lo_c[np.arange(len(row)), ok_lo.argmax(1)]
    hi = hi_c[np.arange(len(row)), ok_hi.argmax(1)]
    bounds = np.searchsorted(slab, np.arange(len(radii) + 1))
    for s_, (b0, b1) in enumerate(zip(bounds[:-1], bounds[1:])):
        f = full[b0:b1]
        if not f.any():
            continue
        k = np.flatnonzero(f)
        if np.any(np.diff(k) != 1):
            raise AssertionError("collection disk rows must be consecutive")
        l_, h_ = lo[b0:b1][k], hi[b0:b1][k]
        c0 = int(l_.min())
        runs = np.stack([l_ - c0, h_ - c0], 1).astype(np.int32).ravel()
        out[s_] = (runs, len(k), int(h_.max()) - c0 + 1, int((h_ - l_ + 1).sum()))
    return out


def collection_mask(vessels, config, events=None):
    """Depth-dependent collection disk and hemoglobin absorption (optics.py:384-415).

    For every slab of prop_um, the vessel projection is convolved with a normalised disk whose
    radius follows the collection cone, 'same' placement as scipy.signal.fftconvolve; the sum
    over slabs, cropped to the volume, gives 10 ** (-total / vres * hemoabs).  The convolution
    is an exact FP64 sum of row runs over prefix sums, every slab in one launch.  GPU-resident
    vessels are read in place; others are read slab by slab through indexing.  Returns float32
    NumPy.  With `events` (a list), a ("collection_mask", (bytes,), (start, end)) HIP event pair
    is appended for the launch; bytes counts the vessel slabs read, the FP64 prefix and partial
    tables written and read once, and the output."""
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
    cone = np.tan(np.arcsin(config.objNA / config.n))
    used, runs_all, weights = [], [], []
    offset = 0

    def radius_at(start):  # the slab's collection radius in voxels
        depth = (start / pc + 0.5) * config.prop_um
        return r * cone * (config.depth_um + config.volume_um[2] / 2 - depth)

    starts = [st for st in range(0, Zv, pc) if radius_at(st) > 0]
    disks = _disk_runs_many(X, Y, [radius_at(st) for st in starts])
    for start, disk in zip(starts, disks):
        if disk is None:
            continue
        runs, K0, K1, count = disk
        used.append([start, min(start + pc, Zv), offset, K0, (K0 - 1) // 2, (K1 - 1) // 2])
        runs_all.append(runs)
        weights.append(float(np.float32(1.0) / np.float32(count)))
        offset += len(runs)
    src = _device_source(a)
    if src is not None:
        vol, eb = src
    elif used:
        # stage the slabs that contribute, read through indexing, contiguously on the device
        read = _reader(a, dev, raw=True)
        parts = [read(slice(None), slice(None), slice(z0, z1)) for z0, z1, *_ in used]
        if len({t.dtype for t in parts}) > 1:
            parts = [t.to(torch.float32) for t in parts]
        vol = torch.cat(parts, dim=2)
        eb = _elem_bytes(vol)
        z = 0
        for u, t in zip(used, parts):
            u[0], u[1] = z, z + int(t.shape[2])
            z += int(t.shape[2])
    ns = len(used)
    lib, stream = nat.lib(), nat.stream_handle()
    out = torch.empty((Xo, Yo), dtype=torch.float32, device=dev)
    if ns:
        slabs = torch.from_numpy(np.asarray(used, np.int32).ravel()).to(dev)
        wts = torch.tensor(weights, dtype=torch.float64).to(dev)
        runs = torch.from_numpy(np.concatenate(runs_all).astype(np.int32)).to(dev)
        ws = torch.empty(int(lib.maomi_collection_mask_workspace(X, Y, Xo, Yo, ns)),
                         dtype=torch.float64, device=dev)
        nbytes = (X * Y * sum(u[1] - u[0] for u in used) * eb
                  + ns * (8 * X * Y + 16 * X * (Y + 1) + 16 * Xo * Yo) + 4 * Xo * Yo)
        ev = _record(events, "collection_mask", (nbytes,))
        nat.check(lib.maomi_collection_mask(nat.ptr(vol), eb, X, Y, int(vol.shape[2]),
                                            vol.stride(0), vol.stride(1), nat.ptr(slabs),
                                            nat.ptr(wts), nat.ptr(runs), ns, ox, oy, Xo, Yo,
                                            float(np.float32(r)),
                                            float(np.float32(config.hemoabs)), nat.ptr(ws),
                                            nat.ptr(out), stream), "maomi_collection_mask")
        if ev is not None:
            ev[1].record()
    else:
        nat.check(lib.maomi_collection_mask(None, 1, X, Y, 1, Y, 1, None, None, None, 0, ox, oy,
                                            Xo, Yo, float(np.float32(r)),
                                            float(np.float32(config.hemoabs)), None,
                                            nat.ptr(out), stream), "maomi_collection_mask")
    return out.cpu().numpy()
