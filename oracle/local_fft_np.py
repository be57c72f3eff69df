"""CPU restatement of MaoMi's local-PSF FFT plan -- TEST INFRASTRUCTURE ONLY.

``local_fft`` follows maomi/rana/_fft.py:8-99 (tile grouping by weight support, zero-padded
rFFT convolution or correlation per tile, plane projection by a sum over z, weighted scatter
into the output) in FP64 with scipy.fft, one tile at a time.  ``local_direct`` evaluates the
same operator as a direct sum over kernel offsets; tests/test_oracle_local_fft.py pins the
tiled restatement to it on small cases of every mode.  ``axis_weights`` restates the smooth
node weights of maomi/rana/operator.py:10-17.

The operator, for kernel nodes (i, j) with XY weights W_ij(x, y) = wx[i, x] wy[j, y]:
  detection   out[p] = sum_ij sum_q W_ij(q) s[q] k_ij[c + p - q],        c = K // 2
  excitation  out[p] = sum_ij W_ij(p) sum_q s[q] k_ij[K // 2 + q - p]     (correlation)
  native      out[p] = sum_ij W_ij(p) sum_q s[q] k_ij[c + p - q]
with zero source outside the volume.  In a plane projection the z index of the kernel equals
the source z and the output sums over z; otherwise z is convolved like x and y.
"""

import numpy as np
from scipy import fft


def axis_weights(nodes, positions):
    """operator.py:10-17: each position gets smoothstep weights on its two bracketing nodes."""
    nodes = np.asarray(nodes, float)
    positions = np.asarray(positions, float)
    k = np.clip(np.searchsorted(nodes, positions, side="right") - 1, 0, len(nodes) - 2)
    t = np.clip((positions - nodes[k]) / (nodes[k + 1] - nodes[k]), 0.0, 1.0)
    t = t * t * (3 - 2 * t)
    w = np.zeros((len(nodes), len(positions)))
    cols = np.arange(len(positions))
    w[k, cols] = 1 - t
    w[k + 1, cols] = t
    return w


def _tiles(shape, kshape, weights, detection, correlate):
    """Per (i, j) with nonzero support: source window, target window, response start and
    the weights tile (rana/_fft.py:22-47)."""
    kxy = np.array(kshape[:2])
    centre = np.array([(k - 1 - k // 2) if correlate else k // 2 for k in kshape])
    wx, wy = weights
    out = []
    for i in range(wx.shape[0]):
        for j in range(wy.shape[0]):
            sx, sy = np.flatnonzero(wx[i]), np.flatnonzero(wy[j])
            if len(sx) == 0 or len(sy) == 0:
                continue
            lo = np.array([sx[0], sy[0]])
            hi = np.array([sx[-1] + 1, sy[-1] + 1])
            w = wx[i, lo[0]:hi[0], None] * wy[j, None, lo[1]:hi[1]]
            if detection:
                s0, s1 = lo, hi
                t0 = np.maximum(0, lo - centre[:2])
                t1 = np.minimum(shape[:2], hi + kxy - 1 - centre[:2])
                r0 = t0 - lo + centre[:2]
            else:
                s0 = np.maximum(0, lo - kxy)
                s1 = np.minimum(shape[:2], hi + kxy)
                t0, t1 = lo, hi
                r0 = lo - s0 + centre[:2]
            out.append(dict(i=i, j=j, s0=s0, s1=s1, t0=t0, t1=t1, r0=r0, rz=int(centre[2]), w=w))
    return out


def local_fft(shape, kernels, weights, source, *, plane, detection, correlate):
    """The tiled FFT plan applied to one source volume (FP64)."""
    shape = tuple(shape)
    source = np.asarray(source, float)
    kernels = np.asarray(kernels, float)
    kshape = kernels.shape[2:]
    axes = (0, 1) if plane else (0, 1, 2)
    out = np.zeros(shape[:2] if plane else shape)
    for t in _tiles(shape, kshape, weights, detection, correlate):
        part = source[t["s0"][0]:t["s1"][0], t["s0"][1]:t["s1"][1]]
        if detection:
            part = part * t["w"][..., None]
        k = kernels[t["i"], t["j"]]
        if correlate:
            k = np.flip(k, axis=(0, 1) if plane else (0, 1, 2))
        n = [fft.next_fast_len(part.shape[d] + kshape[d] - 1) for d in axes]
        prod = fft.rfftn(part, n, axes=axes) * fft.rfftn(k, n, axes=axes)
        if plane:
            prod = prod.sum(axis=2)
        full = fft.irfftn(prod, n, axes=(0, 1) if plane else axes)
        tx, ty = t["t1"] - t["t0"]
        if tx <= 0 or ty <= 0:
            continue
        r = full[t["r0"][0]:t["r0"][0] + tx, t["r0"][1]:t["r0"][1] + ty]
        if not plane:
            r = r[:, :, t["rz"]:t["rz"] + shape[2]]
        if not detection:
            r = r * (t["w"] if plane else t["w"][..., None])
        out[t["t0"][0]:t["t1"][0], t["t0"][1]:t["t1"][1]] += r
    return out


def _shifted(a, d):
    """b[p] = a[p + d] with zero outside, over the leading len(d) axes."""
    b = np.zeros_like(a)
    src, dst = [], []
    for ax, s in enumerate(d):
        n = a.shape[ax]
        lo, hi = max(0, -s), min(n, n - s)
        if lo >= hi:
            return b
        dst.append(slice(lo, hi))
        src.append(slice(lo + s, hi + s))
    b[tuple(dst)] = a[tuple(src)]
    return b


def local_direct(shape, kernels, weights, source, *, plane, detection, correlate):
    """The same operator as a direct sum over kernel offsets (brute force, small cases)."""
    shape = tuple(shape)
    source = np.asarray(source, float)
    kernels = np.asarray(kernels, float)
    kshape = kernels.shape[2:]
    wx, wy = weights
    out = np.zeros(shape[:2] if plane else shape)
    for i in range(wx.shape[0]):
        for j in range(wy.shape[0]):
            W = np.outer(wx[i], wy[j])
            if not W.any():
                continue
            s = source * W[..., None] if detection else source
            acc = np.zeros_like(out)
            k = kernels[i, j]
            for u in np.ndindex(*(kshape[:2] if plane else kshape)):
                u = np.array(u)
                # correlation reads s[p + u - K//2], convolution reads s[p - u + K//2]
                d = u - np.array(kshape[:len(u)]) // 2 if correlate else np.array(kshape[:len(u)]) // 2 - u
                sh = _shifted(s, d)
                if plane:
                    acc += (sh * k[u[0], u[1]][None, None, :]).sum(axis=2)
                else:
                    acc += sh * k[tuple(u)]
            out += acc if detection else acc * (W if plane else W[..., None])
    return out
