"""MATLAB-compatible numerics of the optical stage (call surface of maomi/numerics.py).

* resize_xy / resize_xy_region / translate_xy: the reference semantics, in torch
  (plan-time helpers); the hot phase preparation runs maomi_resize_phase.
* resize_plan / resample_matrix: the host-built tap tables that the HIP kernels
  consume (the reference builds the same tables in NumPy, native.py:27-48).
* fresnel: scaled multi-screen Fresnel propagation with HIP FFTs (maomi_fft2_c64).
* v4_interpolate: MATLAB griddata 'v4' with the FP64 Green build and evaluation
  on the GPU (maomi_v4_green_matrix_f64, maomi_v4_apply_f64) and the FP64 dense
  solve from torch.linalg.
"""

import numpy as np
import torch

from . import _native as nat


def _cubic(x):
    """Keys cubic kernel, a = -0.5."""
    a = np.abs(x)
    inner = ((1.5 * a - 2.5) * a * a + 1) * (a <= 1)
    outer = (((-0.5 * a + 2.5) * a - 4) * a + 2) * ((a > 1) & (a < 2))
    return inner + outer


def resize_plan(n, scale, start=0, count=None):
    """Tap indices and weights of imresize bicubic along one axis of length n.

    Output samples start..start+count-1 sit at centres (i + 1/2)/scale - 1/2;
    downscaling widens the kernel by 1/scale (antialiasing); indices reflect
    symmetrically at the borders; weights are normalised per output sample.
    """
    m = int(np.ceil(n * scale)) if count is None else count
    centres = (np.arange(start, start + m) + 0.5) / scale - 0.5
    shrink = min(scale, 1)
    width = 4 / shrink
    idx = np.floor(centres - width / 2).astype(int)[:, None] + np.arange(int(np.ceil(width)) + 2)[None]
    w = _cubic((centres[:, None] - idx) * shrink) * shrink
    w /= w.sum(1, keepdims=True)
    idx %= 2 * n
    idx = np.where(idx < n, idx, 2 * n - 1 - idx)
    return idx, w


def resize_xy(a, scale):
    """imresize(a, scale, 'bicubic') over the first two axes (numerics.py:15-39)."""
    out = a
    for axis in (0, 1):
        idx, w = resize_plan(out.shape[axis], scale)
        ids = torch.as_tensor(idx, device=out.device)
        wt = torch.as_tensor(w, device=out.device, dtype=out.real.dtype)
        moved = out.movedim(axis, 0)
        g = moved[ids]
        out = (g * wt.reshape(wt.shape + (1,) * (moved.ndim - 1))).sum(1).movedim(0, axis)
    return out


def resize_xy_region(a, scale, starts, shape, *, device=None):
    """Region [starts, starts + shape) of resize_xy(a, scale), reading only its halo."""
    plans = []
    for axis in (0, 1):
        idx, w = resize_plan(a.shape[axis], scale, starts[axis], shape[axis])
        plans.append((idx, w, int(idx.min()), int(idx.max()) + 1))
    out = a[plans[0][2]:plans[0][3], plans[1][2]:plans[1][3]].to(device or a.device)
    for axis, (idx, w, lo, _) in enumerate(plans):
        moved = out.movedim(axis, 0)
        ids = torch.as_tensor(idx - lo, device=out.device)
        wt = torch.as_tensor(w, device=out.device, dtype=out.real.dtype)
        acc = None
        for j in range(idx.shape[1]):
            if not np.any(w[:, j]):
                continue
            term = moved.index_select(0, ids[:, j]) * wt[:, j].reshape((-1,) + (1,) * (moved.ndim - 1))
            acc = term if acc is None else acc + term
        out = acc.movedim(0, axis)
    return out


def translate_xy(a, shift):
    """imtranslate(a, [shift shift]) with bilinear weights and zero fill: out[i, j] = a(i - s, j - s)."""
    k = int(np.floor(shift))
    f = float(shift - k)

    def shifted(t, axis, s):
        n = t.shape[axis]
        res = torch.zeros_like(t)
        if abs(s) >= n:
            return res
        if s >= 0:
            res.narrow(axis, s, n - s).copy_(t.narrow(axis, 0, n - s))
        else:
            res.narrow(axis, 0, n + s).copy_(t.narrow(axis, -s, n + s))
        return res

    tmp = (1 - f) * shifted(a, 0, k) + f * shifted(a, 0, k + 1)
    return (1 - f) * shifted(tmp, 1, k) + f * shifted(tmp, 1, k + 1)


def resample_matrix(n, ss):
    """translate(ss/2 - 1/2) followed by imresize(1/ss), as a dense (ceil(n/ss), n) matrix."""
    scale = 1 / ss
    shift = ss / 2 - 0.5
    idx, w = resize_plan(n, scale)
    rows = np.broadcast_to(np.arange(idx.shape[0])[:, None], idx.shape)
    pos = idx - shift
    base = np.floor(pos).astype(np.int64)
    frac = pos - base
    mat = np.zeros((idx.shape[0], n))
    # the linear translation's two taps per resize tap, accumulated in tap order
    cols = np.stack([base, base + 1], -1).reshape(idx.shape[0], -1)
    vals = np.stack([w * (1 - frac), w * frac], -1).reshape(idx.shape[0], -1)
    r = np.repeat(rows, 2, axis=1)
    ok = (cols >= 0) & (cols < n)
    np.add.at(mat, (r[ok], cols[ok]), vals[ok])
    return mat


_TW = {}


def twiddles(n, device):
    """Device table W_n^j (cached per size and device)."""
    key = (int(n), str(device))
    if key not in _TW:
        t = torch.empty(n, dtype=torch.complex64, device=device)
        nat.check(nat.lib().maomi_twiddles(nat.ptr(t), n, nat.stream_handle(device)), "twiddles")
        _TW[key] = t
    return _TW[key]


def fft2(u, inverse=False):
    """Unnormalised 2-D DFT over the last two (equal) axes of a complex64 CUDA tensor."""
    if u.dtype != torch.complex64 or not u.is_cuda:
        raise ValueError("fft2 needs a complex64 tensor on a ROCm device")
    h, w = u.shape[-2:]
    if h != w:
        raise ValueError("fft2 needs square fields")
    out = u.contiguous().clone()
    batch = out.numel() // (h * w)
    nat.check(nat.lib().maomi_fft2_c64(nat.ptr(out), batch, h, int(inverse),
                                       nat.ptr(twiddles(h, u.device)), nat.stream_handle(u.device)),
              "fft2")
    return out


def fresnel(u, wavelength, dx, z, phi, nidx, *, crop=None, save_from=0, return_all=False):
    """Scaled multi-screen Fresnel propagation (numerics.py:156-210) on the GPU.

    Square even fields; shift-free FFT order (exact for even sizes).  u (..., X, Y);
    phi (..., X, Y, S) with S = len(z) or len(z) - 1.
    """
    dx, z = np.asarray(dx, float).ravel(), np.asarray(z, float).ravel()
    if len(z) < 2 or len(dx) != len(z) or np.any(np.diff(z) <= 0):
        raise ValueError("Need increasing z and one dx per plane")
    if phi.shape[-1] not in (len(z), len(z) - 1):
        raise ValueError("Wrong number of phase screens")
    h, w = u.shape[-2:]
    if h != w or h % 2:
        raise ValueError("fresnel on the GPU needs square fields of even size")
    rdt = u.real.dtype
    x = torch.arange(-h / 2, h / 2, device=u.device, dtype=rdt)[:, None]
    y = torch.arange(-w / 2, w / 2, device=u.device, dtype=rdt)[None, :]
    lam = float(wavelength) / float(nidx)
    k = 2 * np.pi / lam
    dz = np.diff(z)
    sc = dx[1:] / dx[:-1]
    u = u * torch.exp(1j * k * ((x * dx[0]) ** 2 + (y * dx[0]) ** 2) * (1 - sc[0]) / (2 * dz[0])) * phi[..., 0]
    saved = []

    def keep(i, f):
        if return_all and i >= save_from:
            saved.append(f[..., crop[0], crop[1]].clone() if crop else f.clone())

    keep(0, u)
    cache = {}
    for i in range(len(dz)):
        key = (float(dz[i]), float(dx[i]), float(sc[i]))
        if key not in cache:
            q = torch.exp(-1j * np.pi * lam * dz[i] * ((x / (h * dx[i])) ** 2 + (y / (w * dx[i])) ** 2) / sc[i])
            cache[key] = torch.fft.ifftshift(q) / (h * w)
        u = fft2(cache[key] * fft2(u / sc[i]), inverse=True)
        if i + 1 < phi.shape[-1]:
            u = u * phi[..., i + 1]
        keep(i + 1, u)
    out = u * torch.exp(1j * k / 2 * (sc[-1] - 1) / (sc[-1] * dz[-1]) * ((x * dx[-1]) ** 2 + (y * dx[-1]) ** 2))
    return (out, torch.stack(saved, dim=-1)) if return_all else out


def _green_cpu(x, y):
    d = np.sqrt(((x[:, None, :] - y[None, :, :]) ** 2).sum(-1))
    with np.errstate(divide="ignore", invalid="ignore"):
        g = d * d * (np.log(d) - 1)
    g[d == 0] = 0
    return g


def v4_interpolate(points, values, queries, chunk=16384, *, backend="auto", device=None):
    """MATLAB griddata(..., 'v4') in FP64 with one solve shared by all value columns.

    backend "auto" and "cuda" evaluate on the GPU (Green matrix, dense FP64 solve,
    Green apply); "cpu" keeps the reference's NumPy evaluator for GPU-less callers.
    """
    if backend not in ("auto", "cpu", "cuda"):
        raise ValueError("backend must be auto, cpu or cuda")
    if not isinstance(chunk, (int, np.integer)) or chunk < 1:
        raise ValueError("chunk must be a positive integer")
    p, v, q = np.asarray(points, float), np.asarray(values, float), np.asarray(queries, float)
    if p.ndim != 2 or p.shape[1] != 2 or not len(p):
        raise ValueError("points must be a nonempty (N, 2) array")
    if v.ndim not in (1, 2) or len(v) != len(p) or (v.ndim == 2 and v.shape[1] == 0):
        raise ValueError("values must have shape (N,) or (N, C) with C > 0")
    if q.ndim < 1 or q.shape[-1] != 2:
        raise ValueError("queries must have shape (..., 2)")
    shape = q.shape[:-1] + v.shape[1:]
    if len(p) == 1:
        return np.broadcast_to(v[0], shape).copy()
    flat = q.reshape(-1, 2)
    if not len(flat):
        return np.empty(shape, dtype=np.float64)
    if backend == "cpu":
        from scipy.spatial.distance import cdist
        coef = np.linalg.solve(_green_cpu(p, p), v)
        res = np.empty((len(flat),) + v.shape[1:])
        step = min(chunk, max(1, 8_388_608 // len(p)))
        for s in range(0, len(flat), step):
            d2 = cdist(flat[s:s + step], p, "sqeuclidean")
            with np.errstate(divide="ignore", invalid="ignore"):
                g = d2 * (0.5 * np.log(d2) - 1)
            g[d2 == 0] = 0
            res[s:s + step] = g @ coef
        return res.reshape(shape)
    if not torch.cuda.is_available():
        raise RuntimeError("v4_interpolate backend 'cuda' needs a ROCm GPU")
    dev = torch.device(device or "cuda")
    out = v4_gpu(p, v.reshape(len(p), -1), flat, dev, chunk=chunk)
    return out.cpu().numpy().reshape(shape)


def lu_library(on_default_stream):
    """The linear-algebra library for dense_solve_f64's LU.  MAGMA's single-matrix getrf/getrs
    (26 ms against 43 ms for hipSOLVER at P = 4225 on MI355X) run on MAGMA's own queues, which
    are ordered with the device's default stream only; on any other current stream the solve
    could read G before the kernels queued there have written it.  There hipSOLVER, which runs
    on torch's current stream, is used instead."""
    return "magma" if on_default_stream else "cusolver"


def dense_solve_f64(G, v):
    """FP64 LU solve of the v4 system on the device (the library chosen by lu_library)."""
    dev = G.device
    if not G.is_cuda:
        return torch.linalg.solve(G, v).contiguous()
    lib = lu_library(torch.cuda.current_stream(dev) == torch.cuda.default_stream(dev))
    prev = torch.backends.cuda.preferred_linalg_library()
    try:
        torch.backends.cuda.preferred_linalg_library(lib)
        return torch.linalg.solve(G, v).contiguous()
    finally:
        torch.backends.cuda.preferred_linalg_library(prev)


def cholesky_f64_(A, sync=True):
    """In-place FP64 Cholesky of a row-major device view A (m x m, unit column stride) with
    maomi_cholesky_f64: the lower triangle receives L, the upper is left as it was.  Returns
    0, or the 1-based column of the first non-positive pivot (one device-to-host read); with
    sync=False the device int32 tensor that holds it, without reading it."""
    m = A.shape[0]
    if A.dtype != torch.float64 or not A.is_cuda or A.shape != (m, m) or A.stride(1) != 1:
        raise ValueError("cholesky_f64_ needs a square float64 device view with unit column stride")
    lib = nat.lib()
    work = torch.empty(int(lib.maomi_cholesky_f64_workspace()) // 8, dtype=torch.float64,
                       device=A.device)
    info = torch.empty(1, dtype=torch.int32, device=A.device)
    nat.check(lib.maomi_cholesky_f64(nat.ptr(A), A.stride(0), m, nat.ptr(work), nat.ptr(info),
                                     nat.stream_handle(A.device)), "cholesky")
    return int(info) if sync else info


def _reflectors(points):
    """Householder vectors of the QR of P = [1 x y] (host FP64, n x 3): V (n x 3, unit
    columns, H_k = I - 2 v_k v_k^T) and T (3 x 3, upper triangular) with
    H1 H2 H3 = I - V T V^T (compact WY form).  None when P has rank below 3."""
    p = np.asarray(points, np.float64)
    n = len(p)
    Pm = np.stack([np.ones(n), p[:, 0], p[:, 1]], 1)
    scale = float(np.abs(Pm).max())
    V = np.zeros((n, 3))
    for k in range(3):
        x = Pm[k:, k]
        nx = float(np.linalg.norm(x))
        alpha = -nx if x[0] >= 0 else nx
        if nx == 0.0 or abs(alpha) <= 1e-10 * scale * (n ** 0.5):
            return None
        u = x.copy()
        u[0] -= alpha
        nu = float(np.linalg.norm(u))
        if nu == 0.0:
            return None
        V[k:, k] = u / nu
        Pm -= 2.0 * np.outer(V[:, k], V[:, k] @ Pm)
    VtV = V.T @ V
    T = np.zeros((3, 3))
    for i in range(3):
        T[i, i] = 2.0
        if i:
            T[:i, i] = -2.0 * (T[:i, :i] @ VtV[:i, i])
    return V, T


class V4Factor:
    """The part of the v4 solve that depends on the points only (see v4_factor)."""

    def __init__(self, G, G0, V, T, L, A11, A12, info):
        self.G, self.G0, self.V, self.T, self.L = G, G0, V, T, L
        self.A11, self.A12, self.info = A11, A12, info

    def tensors(self):
        """The tensors it owns (L, A11 and A12 are views of G)."""
        return [t for t in (self.G, self.G0, self.V, self.T, self.info)
                if isinstance(t, torch.Tensor)]


def v4_factor(points, G, *, keep_copy=True, reflectors=None):
    """Point-only part of the v4 solve G w = v (FP64, on G's device), through its
    positive-definite part.

    G_ij = r^2 (ln r - 1) is conditionally positive definite of order 2: on vectors orthogonal
    to the polynomials 1, x, y it is positive definite (the r^2 part vanishes there, the thin-
    plate r^2 ln r part is positive).  With Q = H1 H2 H3 the Householder reflectors of the
    QR of P = [1 x y], A = Q^T G Q has A22 = A[3:, 3:] positive definite.  This computes A
    (G is overwritten) and the Cholesky factor L of A22 (maomi_cholesky_f64 on the device).
    The values enter only in v4_finish.
    keep_copy keeps G's original for the LU fallback; reflectors = (V, T) as device tensors, or
    False for rank below 3, from _reflectors computed earlier.  Returns None when P = [1 x y] has rank
    below 3 (points on one line, repeated points); no host synchronisation after the
    reflectors are built (they are built on the host from the points)."""
    n = G.shape[0]
    G0 = G.clone() if keep_copy else None
    if n <= 3:
        return None
    if reflectors is None:
        if isinstance(points, torch.Tensor):
            points = points.detach().cpu().numpy()
        ref = _reflectors(points)
        reflectors = False if ref is None else tuple(
            torch.as_tensor(a, dtype=G.dtype, device=G.device) for a in ref)
    if reflectors is False:
        return V4Factor(None, G0, None, None, None, None, None, None)
    V, T = reflectors
    # Q^T G Q = G - (N V^T + V N^T) with M = G V, N = M T - V S / 2, S = T^T V^T M T: one pass
    # over G for M and one symmetric rank-6 update.
    M = G @ V
    Sm = T.T @ (V.T @ M) @ T
    N = M @ T - 0.5 * (V @ Sm)
    G.addmm_(torch.cat([N, V], 1), torch.cat([V, N], 1).T, alpha=-1.0)
    A11, A12, A22 = G[:3, :3], G[:3, 3:], G[3:, 3:]
    if G.is_cuda:
        info = cholesky_f64_(A22, sync=False)
        L = A22  # lower triangle; the upper keeps A22's own (symmetric) values
    else:
        L, info = torch.linalg.cholesky_ex(A22)
    return V4Factor(G, G0, V, T, L, A11, A12, info)


def v4_finish(f, v):
    """w with G w = v from a v4_factor result; None when the positive-definite route does not
    apply (the caller then solves by LU).  One host read of the factorisation's pivot report.
    On the device the whole finish (Q^T v, the triangular sweeps, the 3 x 3 Schur complement and
    Q x) is maomi_v4_finish_f64."""
    if f is None or f.V is None or int(f.info) != 0:
        return None
    v = v.reshape(f.V.shape[0], -1)
    if f.G.is_cuda:
        return _v4_finish_device(f, v)
    V, T = f.V, f.T
    y = v - V @ (T.T @ (V.T @ v))  # Q^T v
    X = torch.cholesky_solve(torch.cat([f.A12.T, y[3:]], 1), f.L)  # A22^-1 [A12^T y2]
    Z, z = X[:, :3], X[:, 3:]
    a = torch.linalg.solve(f.A11 - f.A12 @ Z, y[:3] - f.A12 @ z)  # 3 x 3 Schur complement
    x = torch.cat([a, z - Z @ a], 0)
    return (x - V @ (T @ (V.T @ x))).contiguous()  # Q x


def _v4_finish_device(f, v):
    """maomi_v4_finish_f64 on the factored system held in f.G (value columns in chunks of at
    most MAOMI_V4_FINISH_MAX_C)."""
    lib = nat.lib()
    G, n = f.G, f.G.shape[0]
    if G.stride(1) != 1 or G.dtype != torch.float64:
        raise ValueError("the factored v4 system must be float64 with unit column stride")
    V, T = f.V.contiguous(), f.T.contiguous()
    v = v.to(torch.float64).contiguous()
    C = v.shape[1]
    w = torch.empty_like(v)
    stream = nat.stream_handle(G.device)
    for c0 in range(0, C, _V4_FINISH_MAX_C):
        vc = v[:, c0:c0 + _V4_FINISH_MAX_C].contiguous()
        k = vc.shape[1]
        wc = w if k == C else torch.empty_like(vc)
        work = torch.empty(int(lib.maomi_v4_finish_f64_workspace(n, k)) // 8,
                           dtype=torch.float64, device=G.device)
        nat.check(lib.maomi_v4_finish_f64(nat.ptr(G), G.stride(0), n, nat.ptr(V), nat.ptr(T),
                                          nat.ptr(vc), k, nat.ptr(wc), nat.ptr(work), stream),
                  "v4 finish")
        if k != C:
            w[:, c0:c0 + k] = wc
    return w


_V4_FINISH_MAX_C = 13  # MAOMI_V4_FINISH_MAX_C


def v4_solve(points, G, v):
    """Solve the v4 system G w = v through its positive-definite part (v4_factor, v4_finish):
    mathematically the solution of numerics.py:129's np.linalg.solve.  Returns None where that
    route does not apply (the caller falls back to LU).  G is modified."""
    n = G.shape[0]
    if n <= 3:
        return dense_solve_f64(G, v.reshape(n, -1))
    return v4_finish(v4_factor(points, G, keep_copy=False), v)


def v4_coefficients(points, G, v, factor=None):
    """v4 weights: the positive-definite route, LU where it does not apply.  `factor` is a
    v4_factor result for these points computed earlier (G is then not used)."""
    if factor is None:
        if G.shape[0] <= 3:
            return dense_solve_f64(G, v.reshape(G.shape[0], -1))
        factor = v4_factor(points, G)
    w = v4_finish(factor, v)
    if w is None:
        return dense_solve_f64(factor.G0, v.reshape(factor.G0.shape[0], -1))
    return w


def _upload(a, device):
    """Host array to the device through pinned memory, without blocking the host."""
    return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(device, non_blocking=True)


def v4_prepare(points, device):
    """Green matrix and the point-only part of the v4 solve (numerics.v4_factor), queued on
    the current stream with no host synchronisation.  Returns (points, G, factor)."""
    pts = np.ascontiguousarray(points, np.float64)
    P = len(pts)
    p = _upload(pts, device)
    ref = _reflectors(pts) if P > 3 else None
    ref = False if ref is None else tuple(_upload(a, device) for a in ref)
    G = torch.empty((P, P), dtype=torch.float64, device=device)
    nat.check(nat.lib().maomi_v4_green_matrix_f64(nat.ptr(p), P, nat.ptr(G),
                                                  nat.stream_handle(device)), "green matrix")
    factor = v4_factor(p, G, reflectors=ref) if P > 3 else None
    return p, G, factor


class V4Ahead:
    """The point-only part of the v4 solve (v4_prepare) queued on a side stream (default
    priority), ordered after the work already queued on the current stream when start() is
    called; take() orders the current stream after it."""

    def __init__(self, device):
        self.device, self.side, self.ready = device, None, None

    def start(self, points):
        main = torch.cuda.current_stream(self.device)
        after = torch.cuda.Event()
        after.record(main)
        if self.side is None:
            self.side = torch.cuda.Stream(self.device)
        self.side.wait_event(after)
        with torch.cuda.stream(self.side):
            prep = v4_prepare(points, self.device)
            done = torch.cuda.Event()
            done.record(self.side)
        self.ready = (prep, done)

    def take(self):
        """The prepared (points, G, factor), with the current stream ordered after it and its
        tensors marked as used there; None if nothing was started."""
        if self.ready is None:
            return None
        prep, done = self.ready
        self.ready = None
        main = torch.cuda.current_stream(self.device)
        main.wait_event(done)
        p, G, factor = prep
        for t in [p, G] + (factor.tensors() if factor is not None else []):
            t.record_stream(main)
        return prep


def v4_apply(p, coef, q, out, device):
    """out (Q, C) = sum_i G(|q - p_i|) coef[i] on the device, the point sum split over block
    rows so that small query sets still fill the chip (maomi_v4_apply_split_f64)."""
    lib = nat.lib()
    P, C, Q = p.shape[0], coef.shape[1], q.shape[0]
    rows = int(lib.maomi_v4_apply_rows(P, Q))
    part = torch.empty(rows * Q * C if rows > 1 else 0, dtype=torch.float64, device=device)
    nat.check(lib.maomi_v4_apply_split_f64(nat.ptr(p), nat.ptr(coef), nat.ptr(q), nat.ptr(out),
                                           P, Q, C, nat.ptr(part) if rows > 1 else None, rows,
                                           nat.stream_handle(device)), "v4 apply")
    return out


# The table of G for grid evaluation holds U^2 doubles for U distinct axis distances.
_GRID_MAX_U = 2048
_GRID_MAX_NY = 128


def point_grid(points):
    """(xs, ys, order) when the points are a complete tensor grid of their distinct x and y
    values, point order[j * len(ys) + k] sitting at (xs[j], ys[k]); else None."""
    pts = np.asarray(points, np.float64)
    if pts.ndim != 2 or pts.shape[1] != 2 or len(pts) < 2:
        return None
    xs, jx = np.unique(pts[:, 0], return_inverse=True)
    ys, ky = np.unique(pts[:, 1], return_inverse=True)
    if len(xs) * len(ys) != len(pts):
        return None
    cell = jx.ravel() * len(ys) + ky.ravel()
    if np.unique(cell).size != len(pts):
        return None
    order = np.empty(len(pts), np.int64)
    order[cell] = np.arange(len(pts))
    return xs, ys, order


def grid_tables(grid, qx, qy):
    """Distinct axis distances du and the index tables of maomi_v4_apply_grid_f64 for points
    on `grid` (point_grid) and queries (qx[a], qy[b]); None past the table's bounds."""
    xs, ys, _ = grid
    if len(ys) > _GRID_MAX_NY:
        return None
    dx = np.abs(np.asarray(qx, np.float64)[:, None] - xs[None, :])
    dy = np.abs(np.asarray(qy, np.float64)[:, None] - ys[None, :])
    du, inv = np.unique(np.concatenate([dx.ravel(), dy.ravel()]), return_inverse=True)
    if len(du) > _GRID_MAX_U:
        return None
    inv = inv.ravel().astype(np.int32)
    return du, inv[:dx.size], inv[dx.size:]


def v4_apply_grid(grid, coef, qx, qy, out, device):
    """out (len(qx) len(qy), C) = sum_i G(|q - p_i|) coef[i] for points on `grid` (point_grid,
    coef in the points' order) and queries q = (qx[a], qy[b]) at row a len(qy) + b, with G
    from a table (maomi_v4_apply_grid_f64).  Returns out, or None when the grid is outside the
    table's bounds (the caller then uses v4_apply)."""
    tabs = grid_tables(grid, qx, qy)
    if tabs is None:
        return None
    du, ix, iy = tabs
    xs, ys, order = grid
    nx, ny, mx, my = len(xs), len(ys), len(qx), len(qy)
    C, U = coef.shape[1], len(du)
    lib = nat.lib()
    cg = coef.index_select(0, _upload(order, device)).contiguous()
    d_du, d_ix, d_iy = (_upload(a, device) for a in (du, ix, iy))
    table = torch.empty(U * U, dtype=torch.float64, device=device)
    rows = min(int(lib.maomi_v4_apply_rows(nx * ny, mx * my)), nx)
    part = torch.empty(rows * mx * my * C if rows > 1 else 0, dtype=torch.float64, device=device)
    nat.check(lib.maomi_v4_apply_grid_f64(nat.ptr(d_du), U, nat.ptr(table), nat.ptr(d_ix),
                                          nat.ptr(d_iy), nat.ptr(cg), nat.ptr(out), nx, ny, mx,
                                          my, C, nat.ptr(part) if rows > 1 else None, rows,
                                          nat.stream_handle(device)), "v4 apply on a grid")
    return out


def v4_gpu(points, values, flat_queries, device, chunk=1 << 20, prepared=None):
    """Device FP64 v4: returns a (Q, C) float64 tensor on `device`.  prepared: a v4_prepare
    result for these points (e.g. from V4Ahead.take())."""
    p, G, factor = prepared if prepared is not None else v4_prepare(points, device)
    v = torch.as_tensor(values, dtype=torch.float64, device=device).reshape(len(p), -1).contiguous()
    q = torch.as_tensor(np.ascontiguousarray(flat_queries, np.float64), device=device)
    P, C = v.shape
    coef = v4_coefficients(p, G, v, factor=factor)
    out = torch.empty((q.shape[0], C), dtype=torch.float64, device=device)
    for s in range(0, q.shape[0], chunk):
        v4_apply(p, coef, q[s:s + chunk], out[s:s + chunk], device)
    return out
