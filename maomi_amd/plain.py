"""Plain optical stage on one MI355X: phase prep, beam prep, all sites in one launch, v4 masks.

Timing boundary as the reference's run_fixed (maomi/fixed.py:337-378):
plan (phase + beam preparation) + propagation (all sites) + aggregation (mask
assembly).  Every numerical step runs in libmaomi_amd.so; torch provides
device memory, streams and the FP64 dense solve of the v4 system.
"""

import ctypes
import time

import numpy as np
import torch

from . import _native as nat
from .numerics import V4Ahead, v4_prepare
from .workloads import PlainConfig, beam_parameters, site_offsets, zernike_weights

_CLASSES = {64: 0, 128: 1, 256: 2, 512: 3}
_SLOT = 512  # table slots per class


def _check_sizes(n):
    if n not in _CLASSES:
        raise ValueError(f"field side {n} unsupported; use one of {sorted(_CLASSES)}")


class SiteTables:
    """Per-class edge windows, 1-D Fresnel factors and twiddles on the device."""

    def __init__(self, sizes, dx, dz, lam_eff, device):
        lib = nat.lib()
        st = nat.stream_handle(device)
        self.edge = torch.zeros(5 * _SLOT, dtype=torch.float32, device=device)
        self.q = torch.zeros(5 * _SLOT, dtype=torch.complex64, device=device)
        self.tw = torch.zeros(5 * _SLOT, dtype=torch.complex64, device=device)
        self.edge2d = {}
        for n in sorted(set(int(s) for s in sizes)):
            _check_sizes(n)
            c = _CLASSES[n]
            e2 = torch.empty((n, n), dtype=torch.float32, device=device)
            e1 = self.edge[c * _SLOT:c * _SLOT + n]
            nat.check(lib.maomi_edge_window(nat.ptr(e1), nat.ptr(e2), n, st), "edge window")
            nat.check(lib.maomi_fresnel_q1d(nat.ptr(self.q[c * _SLOT:]), n, float(dx), float(dz),
                                            float(lam_eff), st), "fresnel q1d")
            nat.check(lib.maomi_twiddles(nat.ptr(self.tw[c * _SLOT:]), n, st), "twiddles")
            self.edge2d[n] = e2
        self.offsets = [c * _SLOT for c in range(5)]


def make_pupil(n, bp, zern, device):
    """generate_ba on an n x n grid (gaussian beam, objective aperture, Zernike screen)."""
    lib = nat.lib()
    st = nat.stream_handle(device)
    u = torch.empty((n, n), dtype=torch.complex64, device=device)
    abb = np.asarray(zern, float) * bp["wvl_um"] * 1e-6 if len(zern) else np.zeros(0)
    nz = int(len(abb)) if np.any(abb != 0) else 0
    zw = torch.as_tensor(abb.astype(np.float32), device=device) if nz else None
    zmap = torch.empty((n, n), dtype=torch.float32, device=device) if nz else None
    dwork = torch.zeros(1, dtype=torch.float64, device=device)
    nat.check(lib.maomi_pupil_gaussian(nat.ptr(u), n, bp["d1"], bp["rad"], bp["objrad"], bp["k"],
                                       bp["fl"], nat.ptr(zw), nz, nat.ptr(zmap), nat.ptr(dwork),
                                       st), "pupil")
    return u, (zw, zmap, dwork)


def base_beam(u, bp, edge2d, tw, device, *, normalize=True):
    """One scaled Fresnel step d1 -> dx to the entry plane, then unit L2 norm (batched.py:43-51)."""
    lib = nat.lib()
    st = nat.stream_handle(device)
    n = u.shape[-1]
    b = u.clone()
    nat.check(lib.maomi_fresnel_step(nat.ptr(b), 1, n, float(bp["d1"]), float(bp["dx"]),
                                     float(bp["dz_focus"]), float(bp["wvl"]), float(bp["nr"]),
                                     nat.ptr(edge2d), nat.ptr(edge2d), nat.ptr(tw), st),
              "fresnel step")
    if normalize:
        dwork = torch.zeros(1, dtype=torch.float64, device=device)
        nat.check(lib.maomi_normalize_l2(nat.ptr(b), b.numel(), nat.ptr(dwork), st), "normalize")
    return b


def base_beams(u, bp, dz_focus, edge2d, tw, out, device):
    """base_beam of pupil u for several entry distances at once: out[b] = unit-norm Fresnel step
    d1 -> dx over dz_focus[b] (batched.py:43-51 per depth), one batched launch sequence."""
    lib = nat.lib()
    st = nat.stream_handle(device)
    n = u.shape[-1]
    B = len(dz_focus)
    if tuple(out.shape) != (B, n, n) or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError("out must be a contiguous complex64 (B, n, n) tensor")
    lam = float(bp["wvl"]) / float(bp["nr"])
    k = 2.0 * np.pi / lam
    sc = float(bp["dx"]) / float(bp["d1"])
    par = np.empty((B, 3), np.float32)
    for b, dz in enumerate(float(z) for z in dz_focus):
        # the float roundings of maomi_fresnel_step's scalar arguments
        par[b] = (np.float32(1.0) / np.float32(2.0 * dz), np.float32(-np.pi * lam * dz),
                  np.float32(k / 2.0 * (sc - 1.0) / (sc * dz)))
    par_d = torch.as_tensor(par, device=device)
    dwork = torch.zeros(B, dtype=torch.float64, device=device)
    out.copy_(u.expand(B, n, n))
    nat.check(lib.maomi_fresnel_steps(nat.ptr(out), B, n, float(bp["d1"]), float(bp["dx"]),
                                      nat.ptr(par_d), float(bp["wvl"]), float(bp["nr"]),
                                      nat.ptr(edge2d), nat.ptr(edge2d), nat.ptr(tw),
                                      nat.ptr(dwork), st), "fresnel steps")
    return out


def v4_masks(points, values, shape, device, prepared=None):
    """MATLAB v4 on the GPU in FP64: solve G c = v, evaluate on pixel centres (numerics.py:93-153).
    prepared: a v4_prepare result for these points."""
    from .numerics import point_grid, v4_apply, v4_apply_grid, v4_coefficients
    v = torch.as_tensor(values, dtype=torch.float64, device=device)
    if v.ndim == 1:
        v = v[:, None]
    P, C = v.shape
    if P == 1:
        return v[0].expand(shape[0], shape[1], C).clone()
    p, G, factor = prepared if prepared is not None else v4_prepare(points, device)
    coef = v4_coefficients(p, G, v, factor=factor)
    out = torch.empty((shape[0] * shape[1], C), dtype=torch.float64, device=device)
    # pixel centres; on a lattice of points G comes from a table of the few axis distances
    qx, qy = np.arange(shape[0]) + 0.5, np.arange(shape[1]) + 0.5
    grid = point_grid(points)
    if grid is None or v4_apply_grid(grid, coef, qx, qy, out, device) is None:
        xs = torch.as_tensor(qx, dtype=torch.float64, device=device)
        ys = torch.as_tensor(qy, dtype=torch.float64, device=device)
        q = torch.stack(torch.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2).contiguous()
        v4_apply(p, coef, q, out, device)
    return out.reshape(shape[0], shape[1], C)


# Screen phases must stay below this magnitude (radians): the sites kernel's sin/cos
# reduction is exact only there.  A physical screen is a few radians.
SCREEN_PHASE_LIMIT = 131072.0


def check_screen_range(slice_max):
    """Reject phase screens outside the sites kernel's sin/cos domain (one small D2H)."""
    m = float(slice_max.max())
    if not m < SCREEN_PHASE_LIMIT:
        raise ValueError(f"phase screen magnitude {m:.6g} rad outside the supported range "
                         f"(< {SCREEN_PHASE_LIMIT:g} rad); check the index volume and n0")


class PlainStage:
    """One plain optical-stage run on `device` (uniform N and S over all sites)."""

    def __init__(self, cfg: PlainConfig, index_volume, *, device="cuda", precision="fp32",
                 grid=0, sites=None, screens="phase"):
        """screens: "phase" reads the FP32 phases and evaluates sin/cos per site and pixel;
        "complex" precomputes exp(i phase) once per volume voxel for FP32 runs and reads
        8-byte screens (identical results; measured 1-2% slower on the 4225-site workload,
        whose sites kernel waits on memory rather than on arithmetic)."""
        if not torch.cuda.is_available():
            raise RuntimeError("PlainStage needs a ROCm GPU (torch.cuda.is_available() is False)")
        _check_sizes(cfg.n)
        if precision not in ("fp32", "mixed"):
            raise ValueError("precision must be fp32 or mixed")
        if screens not in ("complex", "phase"):
            raise ValueError("screens must be complex or phase")
        self.screens = screens
        self.cfg, self.device, self.precision = cfg, torch.device(device), precision
        vol = torch.as_tensor(index_volume, dtype=torch.float32, device=self.device)
        exp = (cfg.padded, cfg.padded, cfg.slices * cfg.voxels_per_slice)
        if tuple(vol.shape) != exp:
            raise ValueError(f"index volume must have shape {exp}, got {tuple(vol.shape)}")
        self.volume = vol.contiguous()
        self.offsets = site_offsets(cfg) if sites is None else np.asarray(sites, np.int64)
        if np.any(self.offsets < 0) or np.any(self.offsets > cfg.fov):
            raise ValueError("site windows must lie inside the padded volume")
        self.grid = int(grid)
        self.block_threads = 0  # 0: the library's default (512)
        self.bp = dict(beam_parameters(cfg), wvl_um=cfg.wavelength_um)
        self.zern = zernike_weights(cfg)
        self.timing = {}
        P = len(self.offsets)
        d = self.device
        self.site_x = torch.as_tensor(self.offsets[:, 0].astype(np.int32), device=d)
        self.site_y = torch.as_tensor(self.offsets[:, 1].astype(np.int32), device=d)
        self.site_n = torch.full((P,), cfg.n, dtype=torch.int32, device=d)
        self.site_s = torch.full((P,), cfg.slices, dtype=torch.int32, device=d)
        self.order = torch.arange(P, dtype=torch.int32, device=d)
        self.base_off = torch.zeros(P, dtype=torch.int64, device=d)
        self.mask_off = torch.arange(P, dtype=torch.int64, device=d) * (cfg.n * cfg.n)

    # -------------------------------------------------------------- stages
    def plan(self):
        cfg, d = self.cfg, self.device
        lib = nat.lib()
        st = nat.stream_handle(d)
        Z = cfg.slices
        H = W = cfg.padded
        self.phase = torch.empty((Z, H, W), dtype=torch.float32, device=d)
        self.phase16 = self.scale = None
        smax = torch.empty(Z, dtype=torch.float32, device=d)
        if self.precision == "mixed":
            self.phase16 = torch.empty((Z, H, W), dtype=torch.float16, device=d)
            self.scale = torch.empty(Z, dtype=torch.float32, device=d)
        nat.check(lib.maomi_phase_prepare(nat.ptr(self.volume), H, W, Z * cfg.voxels_per_slice,
                                          cfg.voxels_per_slice, float(cfg.n0),
                                          float(self.bp["k0dz"]), nat.ptr(self.phase),
                                          nat.ptr(self.phase16), nat.ptr(self.scale),
                                          nat.ptr(smax), st), "phase prepare")
        check_screen_range(smax)
        self.screen = None
        if self.precision == "mixed":
            self.phase = None  # only the float16 screens are read by the sites kernel
        elif self.screens == "complex":
            self.screen = torch.empty((Z, H, W), dtype=torch.complex64, device=d)
            nat.check(lib.maomi_screen_exp(nat.ptr(self.phase), Z * H * W, nat.ptr(self.screen),
                                           st), "screen exp")
        self.tables = SiteTables([cfg.n], self.bp["dx"], self.bp["dz"],
                                 float(self.bp["wvl"]) / float(self.bp["nr"]), d)
        c = _CLASSES[cfg.n]
        tw = self.tables.tw[c * _SLOT:c * _SLOT + cfg.n]
        self.pupil, _ = make_pupil(cfg.n, self.bp, self.zern, d)
        self.base = base_beam(self.pupil, self.bp, self.tables.edge2d[cfg.n], tw, d)
        return self

    def propagate(self, subset=None, psf=None, tail_hook=None, order=None, audit=None,
                  schedule="static"):
        """All sites (or the indices `subset`) in one persistent launch.

        psf = (out, z0, nz, c0, w) also writes |u|^4 of planes z0..z0+nz-1 cropped to
        [c0, c0+w)^2 into out[nz][P][w][w] (surrogate training targets and audits).

        tail_hook: when the sites do not fill whole rounds of the grid, the launch is issued as
        the whole rounds and then the last partial round, and tail_hook() is called in between
        (work it queues on another stream runs on the CUs that the last round leaves idle).

        order: the processing order of the sites (a permutation of 0..P-1; default natural).
        audit: with psf, an in-launch audit (maomi_sites_audit): dict of device tensors
        pred (nz, A, w, w), slot (P,) int32, state (2,) int32, rel (A,) float32, done (P,)
        int32, and max_fail, tol; the first A entries of `order` are the audited sites.
        schedule: "static" (block b takes order entries b, b + grid, ...) or "queue" (blocks
        take the next entry from a work counter).  Sites of one size run faster static (4225
        sites: 103 ms, against 114 ms from the counter).
        """
        cfg, d = self.cfg, self.device
        lib = nat.lib()
        given = order
        if subset is None:
            site_x, site_y, order = self.site_x, self.site_y, self.order
            P = len(self.offsets)
        else:
            sub = np.asarray(subset, np.int64)
            P = len(sub)
            site_x = torch.as_tensor(self.offsets[sub, 0].astype(np.int32), device=d)
            site_y = torch.as_tensor(self.offsets[sub, 1].astype(np.int32), device=d)
            order = torch.arange(P, dtype=torch.int32, device=d)
        if given is not None:
            order = given if torch.is_tensor(given) else torch.as_tensor(
                np.asarray(given, np.int32), device=d)
            if order.dtype != torch.int32 or order.numel() != P:
                raise ValueError("order must hold one int32 entry per site")
        grid = self.grid or int(lib.maomi_sites_default_grid()) or 256
        grid = min(grid, P)
        stride = cfg.n * cfg.n
        if not hasattr(self, "workspace") or self.workspace.numel() < grid * stride:
            self.workspace = torch.empty(grid * stride, dtype=torch.complex64, device=d)
        self.masks = torch.empty((P, cfg.n, cfg.n), dtype=torch.float32, device=d)
        self.energies = torch.empty(P, dtype=torch.float64, device=d)
        desc = nat.SitesDesc()
        desc.bases = self.base.data_ptr()
        desc.base_off = self.base_off.data_ptr()
        # complex screens unless the launch needs phases (focal-plane output, 256 threads)
        cplx = (getattr(self, "screen", None) is not None and psf is None
                and self.block_threads in (0, 512))
        desc.phase = (self.phase.data_ptr() if self.phase is not None and not cplx else None)
        desc.screen = self.screen.data_ptr() if cplx else None
        desc.phase16 = self.phase16.data_ptr() if self.phase16 is not None else None
        desc.slice_scale = self.scale.data_ptr() if self.scale is not None else None
        desc.edge = self.tables.edge.data_ptr()
        desc.q1d = self.tables.q.data_ptr()
        desc.twiddle = self.tables.tw.data_ptr()
        desc.site_x, desc.site_y = site_x.data_ptr(), site_y.data_ptr()
        desc.site_n, desc.site_s = self.site_n.data_ptr(), self.site_s.data_ptr()
        desc.order = order.data_ptr()
        desc.mask, desc.mask_off = self.masks.data_ptr(), self.mask_off.data_ptr()
        desc.energy = self.energies.data_ptr()
        desc.workspace = self.workspace.data_ptr()
        if schedule not in ("static", "queue"):
            raise ValueError("schedule must be 'static' or 'queue'")
        if schedule == "queue" and getattr(self, "queue", None) is None:
            self.queue = torch.zeros(1, dtype=torch.int32, device=d)
        desc.queue = self.queue.data_ptr() if schedule == "queue" else None
        desc.scratch_stride = stride
        desc.P, desc.H, desc.W, desc.Z = P, cfg.padded, cfg.padded, cfg.slices
        desc.grid, desc.uniform_n, desc.block_threads = grid, cfg.n, self.block_threads
        for i in range(5):
            desc.class_edge[i] = desc.class_q[i] = desc.class_tw[i] = self.tables.offsets[i]
        if psf is not None:
            out, z0, nz, c0, w = psf
            if tuple(out.shape) != (nz, P, w, w) or out.dtype != torch.float32:
                raise ValueError("psf output must be float32 (nz, P, w, w)")
            if z0 < 0 or z0 + nz > cfg.slices + 1 or c0 < 0 or c0 + w > cfg.n:
                raise ValueError("psf window outside the propagated planes or field")
            desc.psf_out = out.data_ptr()
            desc.psf_z0, desc.psf_nz, desc.psf_c0, desc.psf_w = z0, nz, c0, w
        if audit is not None:
            if psf is None:
                raise ValueError("an audit needs the focal-plane output (psf=)")
            au = nat.SitesAudit()
            for key in ("pred", "slot", "state", "rel", "done"):
                setattr(au, key, audit[key].data_ptr())
            au.n_audit, au.max_fail, au.tol = (audit["pred"].shape[1], int(audit["max_fail"]),
                                               float(audit["tol"]))
            self._audit = au
            desc.audit = ctypes.addressof(au)
        self._desc = desc
        self._keep = (site_x, site_y, order)
        stream = nat.stream_handle(d)
        P1 = P - P % grid
        if tail_hook is not None and psf is None and audit is None and P > grid and P1 < P:
            desc.P = P1
            nat.check(lib.maomi_persistent_sites(desc, stream), "persistent sites")
            tail_hook()
            desc.P, desc.grid = P - P1, P - P1
            desc.order = order.data_ptr() + 4 * P1
            nat.check(lib.maomi_persistent_sites(desc, stream), "persistent sites, last round")
        else:
            nat.check(lib.maomi_persistent_sites(desc, stream), "persistent sites")
        return self.masks, self.energies

    def propagate_all(self, overlap=True):
        """propagate() for every site; with overlap, the point-only part of the v4 solve
        (Green matrix, projection, Cholesky) is queued on a side stream after the whole rounds
        of sites, beside the last partial round, and aggregate() picks it up."""
        if not overlap:
            return self.propagate()
        if getattr(self, "_v4", None) is None:
            self._v4 = V4Ahead(self.device)
        pts = self.offsets.astype(np.float64)
        return self.propagate(tail_hook=lambda: self._v4.start(pts))

    def aggregate(self):
        cfg = self.cfg
        pts = self.offsets.astype(np.float64)
        prep = self._v4.take() if getattr(self, "_v4", None) is not None else None
        masks = v4_masks(pts, self.energies, (cfg.fov, cfg.fov), self.device, prepared=prep)
        m = masks[..., 0]
        mean = m.mean()
        self.mask_raw, self.mask_mean = m, mean
        self.mask = m / mean
        return self.mask

    def run(self):
        """plan + propagation + aggregation, each bracketed by a device synchronisation."""
        t = time.perf_counter()
        self.plan()
        torch.cuda.synchronize(self.device)
        self.timing["plan"] = time.perf_counter() - t
        t = time.perf_counter()
        self.propagate_all()
        torch.cuda.synchronize(self.device)
        self.timing["propagation"] = time.perf_counter() - t
        t = time.perf_counter()
        self.aggregate()
        torch.cuda.synchronize(self.device)
        self.timing["aggregation"] = time.perf_counter() - t
        self.timing["total"] = sum(self.timing[k] for k in ("plan", "propagation", "aggregation"))
        return self
