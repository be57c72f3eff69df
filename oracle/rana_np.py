"""CPU restatement of MaoMi's RANA acquisition operator -- TEST INFRASTRUCTURE ONLY.

Follows maomi/rana/operator.py:20-170 (node weights, depth weights, static, scan_slab) on top
of oracle.local_fft_np (FP64).  Pinned in tests/test_oracle_local_fft.py by closed forms: with
one kernel at every node the operator is the MATLAB convn of oracle.zstack (static) or its
centre plane (scan_slab), and a depth bank of identical fields equals the single field.
"""

import numpy as np

from .local_fft_np import axis_weights, local_fft


class Operator:
    def __init__(self, kernels_list, x_um, y_um, voxel_size_um, operator, *, depths_um=None,
                 native=False, origin_um=(0.0, 0.0), z_origin_um=0.0):
        self.kernels = [np.asarray(k, float) for k in kernels_list]
        self.x_um, self.y_um = np.asarray(x_um, float), np.asarray(y_um, float)
        self.pitch = tuple(float(v) for v in voxel_size_um)
        self.op, self.native = operator, native
        self.depths = None if depths_um is None else np.asarray(depths_um, float)
        self.origin, self.z0 = np.asarray(origin_um, float), float(z_origin_um)
        self.kshape = self.kernels[0].shape[2:]

    def weights(self, shape):
        c = [o + (np.arange(n) + 0.5) * p for o, n, p in zip(self.origin, shape[:2], self.pitch)]
        return axis_weights(self.x_um, c[0]), axis_weights(self.y_um, c[1])

    def depth_weights(self, z):
        if self.depths is None:
            return np.ones((1, np.size(z)))
        z = np.atleast_1d(z)
        return np.stack([np.interp(z, self.depths, row) for row in np.eye(len(self.kernels))])

    def _project(self, src, k, plane=False):
        det = not self.native and self.op == "detection"
        corr = not self.native and self.op == "excitation"
        return local_fft(src.shape, k, self.weights(src.shape), src, plane=plane, detection=det,
                         correlate=corr)

    def static(self, volume):
        a = np.asarray(volume, float)
        z = self.z0 + (np.arange(a.shape[2]) + 0.5) * self.pitch[2]
        det = not self.native and self.op == "detection"
        out = np.zeros_like(a)
        for k, w in zip(self.kernels, self.depth_weights(z)):
            if not np.any(w):
                continue
            out += self._project(a * w, k) if det else self._project(a, k) * w
        return np.maximum(out, 0.0)

    def scan_slab(self, slab, z_start=0, scan_avg=1):
        a = np.asarray(slab, float)
        Z = a.shape[2]
        c = self.kshape[2] // 2
        z = self.z0 + (z_start + np.arange(Z) + 0.5) * self.pitch[2]
        det = not self.native and self.op == "detection"
        out = np.zeros(a.shape[:2])
        starts = np.arange(0, Z, scan_avg) if scan_avg > 1 else None
        for k, w in zip(self.kernels, self.depth_weights(z if det else [z[c]])):
            if not np.any(w):
                continue
            s = a * w if det else a
            if det:
                ids = 2 * c - np.arange(Z)
                k = np.take(k, np.minimum(ids, Z - 1), axis=-1).copy()
                k[..., ids >= Z] = 0
            if starts is not None:
                k = np.add.reduceat(k, starts, axis=-1)
                s = np.add.reduceat(s, starts, axis=-1)
            p = self._project(s, k, plane=True)
            out += p if det else p * w[0]
        return np.maximum(out, 0.0)
