"""Local-PSF FFT plan on the GPU (the call surface of maomi._local_fft / maomi/rana/_fft.py).

``LocalFFTPlan(shape, kernels, weights, *, plane, detection, correlate, device,
batch_size=8, cache_mb=256)`` applies a field of local kernels to a source volume, with the
reference's constructor, ``__call__``, ``clear``, ``stats`` and ``cache_bytes`` (rana/_fft.py:
8-99).  Tiles are grouped by (source window, FFT lengths) as in the reference.  Per batch of
tiles, three HIP kernels of libmaomi_amd.so do the array work:
  * maomi_lfft_gather builds the zero-padded, weighted tiles straight from the volume;
  * maomi_lfft_product multiplies tile and kernel spectra, summing over z in one pass for a
    plane projection (the reference materialises the product, then sums);
  * maomi_lfft_scatter adds each tile's response window, weighted, into the output
    (FP64 atomics; neighbouring tiles overlap).
FP64 rFFTs go through torch.fft (rocFFT).  Kernel spectra are cached per tile batch under
``cache_mb``, as in the reference.  The plan runs on a GPU only; ``device='cpu'`` raises.

The kernels are uploaded once, as a device stack with a leading field axis; ``apply`` projects
through every field of the stack at once (rana.RanaOperator's depth banks): a non-detection plan
transforms each source tile once for all fields and, for a plane projection with one weight per
field, combines the fields' spectra before a single inverse transform; per-z weights of a field
are applied inside the gather (detection) or the scatter (excitation) kernels.
"""

from collections import OrderedDict, defaultdict

import numpy as np
import torch

from . import _native as nat

_GATHER, _SCATTER = 5, 8  # int32 fields per tile descriptor (maomi_amd.h)


def _fast_len(n):
    """FFT length of a tile axis: scipy.fft.next_fast_len(n), the reference's choice
    (rana/_fft.py:48), so that groups and cache keys are the same."""
    from scipy import fft
    return fft.next_fast_len(int(n))


class LocalFFTPlan:
    def __init__(self, shape, kernels, weights, *, plane, detection, correlate, device,
                 batch_size=8, cache_mb=256):
        if str(device) == "cpu":
            raise ValueError("this LocalFFTPlan runs on the GPU: pass device='cuda[:index]'")
        if not torch.cuda.is_available():
            raise RuntimeError("LocalFFTPlan needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.shape, self.kernels = tuple(int(s) for s in shape), kernels
        self.plane, self.detection, self.correlate = bool(plane), bool(detection), bool(correlate)
        self.device, self.batch_size = str(device), int(batch_size)
        self.cuda = True
        self._dev = torch.device(self.device)
        self.cache, self.cache_bytes = OrderedDict(), 0
        self.cache_limit = int(cache_mb * 1024 ** 2)
        self.stats = dict(kernel_fft_builds=0, cache_hits=0)
        self.dims = (1, 2) if plane else (1, 2, 3)
        # The kernels live on the device once, as a stack with a leading field axis: a depth
        # bank's fields share this plan's tiles (rana.RanaOperator passes its whole stack).
        if isinstance(kernels, torch.Tensor):
            stack = kernels.to(device=self._dev, dtype=torch.float64)
            if stack.ndim == 5:
                stack = stack[None]
        else:
            stack = torch.as_tensor(np.ascontiguousarray(kernels, dtype=np.float64),
                                    device=self._dev)[None]
        if stack.ndim != 6:
            raise ValueError("kernels must have axes field_X, field_Y, kernel_X, kernel_Y, kernel_Z")
        self._kstack = stack
        kshape = tuple(stack.shape[3:])
        if len(self.shape) != 3:
            raise ValueError("shape must be (X, Y, Z)")
        if plane and kshape[2] != self.shape[2]:
            raise ValueError("a plane projection needs kernel depth equal to the source depth")
        self.kshape = kshape
        self.centre = np.array([(k - 1 - k // 2) if correlate else k // 2 for k in kshape])
        wx, wy = (np.asarray(w, float) for w in weights)
        if wx.shape[1] != self.shape[0] or wy.shape[1] != self.shape[1]:
            raise ValueError("weights must cover the X and Y extent of shape")
        self.groups = defaultdict(list)
        packed, pos = [], 0
        kxy = np.array(kshape[:2])
        for i in range(wx.shape[0]):
            for j in range(wy.shape[0]):
                sx, sy = np.flatnonzero(wx[i]), np.flatnonzero(wy[j])
                if not len(sx) or not len(sy):
                    continue
                lo = np.array([sx[0], sy[0]])
                hi = np.array([sx[-1] + 1, sy[-1] + 1])
                w = wx[i, lo[0]:hi[0], None] * wy[j, None, lo[1]:hi[1]]
                if detection:  # weights act on the source window = their support
                    s0, s1 = lo, hi
                    t0 = np.maximum(0, lo - self.centre[:2])
                    t1 = np.minimum(self.shape[:2], hi + kxy - 1 - self.centre[:2])
                    r0 = t0 - lo + self.centre[:2]
                else:  # weights act on the output window = their support
                    s0 = np.maximum(0, lo - kxy)
                    s1 = np.minimum(self.shape[:2], hi + kxy)
                    t0, t1 = lo, hi
                    r0 = lo - s0 + self.centre[:2]
                if (t1 - t0).min() <= 0:
                    continue
                src = tuple(int(v) for v in s1 - s0) + (self.shape[2],)
                axes = (0, 1) if plane else (0, 1, 2)
                lengths = tuple(_fast_len(src[d] + kshape[d] - 1) for d in axes)
                packed.append(w.ravel())
                self.groups[(src, lengths)].append(dict(
                    i=i, j=j, s0=s0, t0=t0, t1=t1, r0=r0, w=pos))
                pos += w.size
        self._weights = torch.as_tensor(np.concatenate(packed) if packed else np.zeros(1),
                                        dtype=torch.float64, device=self._dev)

    def array(self, a):
        if isinstance(a, torch.Tensor):
            return a.to(device=self._dev, dtype=torch.float64)
        return torch.as_tensor(np.array(a, dtype=np.float64, order="C"), device=self._dev)

    def _kernel_fft(self, entries, lengths, field=0):
        key = (field, tuple((e["i"], e["j"]) for e in entries), lengths)
        if key in self.cache:
            self.stats["cache_hits"] += 1
            self.cache.move_to_end(key)
            return self.cache[key]
        ii = torch.as_tensor([e["i"] for e in entries], device=self._dev)
        jj = torch.as_tensor([e["j"] for e in entries], device=self._dev)
        k = self._kstack[field][ii, jj]  # gathered on the device: (B, kx, ky, kz)
        if self.correlate:
            k = torch.flip(k, dims=(1, 2) if self.plane else (1, 2, 3))
        spec = torch.fft.rfftn(k.contiguous(), s=lengths, dim=self.dims).contiguous()
        size = spec.numel() * spec.element_size()
        self.stats["kernel_fft_builds"] += 1
        if size <= self.cache_limit:
            while self.cache and self.cache_bytes + size > self.cache_limit:
                _, old = self.cache.popitem(last=False)
                self.cache_bytes -= old.numel() * old.element_size()
            self.cache[key] = spec
            self.cache_bytes += size
        return spec

    def __call__(self, source):
        return self.apply(source)

    def apply(self, source, zweights=None, fweights=None, fields=None):
        """sum_f of the projection of `source` through field f's kernels (every field of the
        stack; one field for a plain plan).

        zweights (F, Z) device tensor: per-z weights of each field, applied to the source when
        the plan is a detection plan and to the responses otherwise (3-D output).
        fweights (F,) device tensor: one weight per field for a non-detection plane projection;
        the fields' spectra are then combined before one inverse transform.
        fields: the field indices to project through (default: all)."""
        src = self.array(source).contiguous()
        if tuple(src.shape) != self.shape:
            raise ValueError(f"source shape {tuple(src.shape)} differs from the plan's {self.shape}")
        F = self._kstack.shape[0]
        use = list(range(F)) if fields is None else [int(f) for f in fields]
        if zweights is not None and tuple(zweights.shape) != (F, self.shape[2]):
            raise ValueError("zweights must be (fields, Z)")
        if fweights is not None and (self.detection or not self.plane or
                                     tuple(fweights.shape) != (F,)):
            raise ValueError("fweights serve non-detection plane projections, one per field")
        X, Y, Z = self.shape
        out = torch.zeros(self.shape[:2] if self.plane else self.shape, dtype=torch.float64,
                          device=self._dev)
        lib, st = nat.lib(), nat.stream_handle(self._dev)
        shared = not self.detection  # one source transform serves every field
        for (sshape, lengths), group in self.groups.items():
            Lx, Ly = lengths[0], lengths[1]
            Lz = Z if self.plane else lengths[2]
            # temporary batch spectra bounded to about 128 MiB (rana/_fft.py:78-81)
            per_tile = max(1, int(np.prod(lengths)) * 64 * (Z if self.plane else 1))
            batch = max(1, min(self.batch_size, 128 * 1024 ** 2 // per_tile))
            for start in range(0, len(group), batch):
                entries = group[start:start + batch]
                B = len(entries)
                gd = np.empty((B, _GATHER), np.int32)
                sd = np.empty((B, _SCATTER), np.int32)
                for n, e in enumerate(entries):
                    gd[n] = (e["s0"][0], e["s0"][1], sshape[0], sshape[1],
                             e["w"] if self.detection else -1)
                    tx, ty = e["t1"] - e["t0"]
                    rz = 0 if self.plane else int(self.centre[2])
                    sd[n] = (e["t0"][0], e["t0"][1], tx, ty, e["r0"][0], e["r0"][1], rz,
                             -1 if self.detection else e["w"])
                    if (e["r0"][0] + tx > Lx or e["r0"][1] + ty > Ly
                            or (not self.plane and rz + Z > Lz)
                            or e["s0"][0] + sshape[0] > X or e["s0"][1] + sshape[1] > Y):
                        raise AssertionError("tile window outside its padded transform")
                gdev = torch.as_tensor(gd, device=self._dev)
                sdev = torch.as_tensor(sd, device=self._dev)
                Tx, Ty = int(sd[:, 2].max()), int(sd[:, 3].max())
                tiles = torch.empty((B, Lx, Ly, Lz), dtype=torch.float64, device=self._dev)
                fz, zo = (1, 1) if self.plane else (Lz, Z)

                def transform(zw):
                    nat.check(lib.maomi_lfft_gather_z(nat.ptr(src), X, Y, Z, nat.ptr(gdev),
                                                      nat.ptr(self._weights), nat.ptr(zw),
                                                      nat.ptr(tiles), B, Lx, Ly, Lz, st),
                              "local fft gather")
                    return torch.fft.rfftn(tiles, dim=self.dims).contiguous()

                def respond(spec, fields, fw, zw):
                    ks = [self._kernel_fft(entries, lengths, f) for f in fields]
                    kstack = ks[0] if len(ks) == 1 else torch.stack(ks)
                    n = int(np.prod(spec.shape[:3])) if self.plane else spec.numel()
                    prod = torch.empty(spec.shape[:3] if self.plane else spec.shape,
                                       dtype=torch.complex128, device=self._dev)
                    nat.check(lib.maomi_lfft_product_fields(
                        nat.ptr(spec), nat.ptr(kstack), ks[0].numel(), nat.ptr(fw), len(ks),
                        nat.ptr(prod), n, Z if self.plane else 0, st), "local fft product")
                    s_ = (Lx, Ly) if self.plane else lengths
                    full = torch.fft.irfftn(prod, s=s_, dim=(1, 2) if self.plane else (1, 2, 3))
                    full = full.contiguous()
                    nat.check(lib.maomi_lfft_scatter_z(
                        nat.ptr(full), B, Lx, Ly, fz, nat.ptr(sdev), nat.ptr(self._weights),
                        nat.ptr(zw), nat.ptr(out), X, Y, zo, Tx, Ty, st), "local fft scatter")

                if shared:
                    spec = transform(None)
                    if zweights is None:  # every field at once, spectra combined
                        fw = None if fweights is None else fweights[use].contiguous()
                        respond(spec, use, fw, None)
                    else:
                        for f in use:
                            respond(spec, [f], None, zweights[f].contiguous())
                else:
                    for f in use:
                        zw = None if zweights is None else zweights[f].contiguous()
                        respond(transform(zw), [f], None, None)
        return out

    def clear(self):
        self.cache.clear()
        self.cache_bytes = 0
