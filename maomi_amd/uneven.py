"""Uneven plain stage: sites with mixed field sides and depths in ONE persistent launch.

Every site s has its own field side N_s in {128, 256, 512} and slice count S_s
(focal depth d_s = S_s dz, focus on the last plane).  All sites run in one
launch of the mixed-size persistent kernel (maomi_persistent_sites with
uniform_n = 0); blocks take sites from a work counter in the order of a list
sorted by cost (S N^2 log N, largest first) and the launch is captured once in
a HIP graph and replayed.  An autotuner picks the grid size, the tile shape
(256- or 512-thread blocks, whose LDS tiles hold block_threads / min(N1, N2) lines of
a field side N = N1 N2), the ordering policy and
the schedule (work counter or static striding) by timing the captured
candidates on the device.
"""

import ctypes
import time
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _native as nat
from .numerics import V4Ahead
from .plain import (SiteTables, _CLASSES, _SLOT, base_beams, check_screen_range, make_pupil,
                    v4_masks)
from .workloads import PlainConfig, beam_parameters, zernike_weights


@dataclass(frozen=True)
class UnevenConfig:
    base: PlainConfig = PlainConfig(fov=512, n=512, slices=60, sites_per_axis=65, corr_vox=3.0)
    sizes: tuple = (128, 256, 512)
    size_probs: tuple = (0.5, 0.4, 0.1)
    depth_um: tuple = (20.0, 120.0)


CONFIG3 = UnevenConfig()


def draw_sites(ucfg: UnevenConfig):
    """Seeded per-site field sides and slice counts (depth ~ U(lo, hi), S = round(d / dz))."""
    cfg = ucfg.base
    rng = np.random.default_rng(cfg.seed + 2)
    P = cfg.num_sites
    depth = rng.uniform(*ucfg.depth_um, P)
    slices = np.clip(np.round(depth / cfg.dz_um).astype(int), 1, cfg.slices)
    sizes = rng.choice(np.asarray(ucfg.sizes), size=P, p=np.asarray(ucfg.size_probs))
    return sizes.astype(np.int64), slices.astype(np.int64)


def site_windows(cfg: PlainConfig, centres, sizes):
    """Top-left of each centred window in a volume padded by cfg.n (the largest side)."""
    return np.asarray(centres, np.int64) + (cfg.n - np.asarray(sizes)[:, None]) // 2


def bytes_per_site(sizes, slices):
    n2 = np.asarray(sizes, np.float64) ** 2
    return (20 * np.asarray(slices) + 12) * n2


class UnevenStage:
    def __init__(self, cfg: PlainConfig, index_volume, centres, sizes, slices, *, device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("UnevenStage needs a ROCm GPU")
        self.cfg, self.device = cfg, torch.device(device)
        self.centres = np.asarray(centres, np.int64)
        self.sizes = np.asarray(sizes, np.int64)
        self.slices = np.asarray(slices, np.int64)
        P = len(self.centres)
        if not (len(self.sizes) == len(self.slices) == P):
            raise ValueError("one size and one slice count per site")
        for n in np.unique(self.sizes):
            if int(n) not in _CLASSES or n > cfg.n:
                raise ValueError(f"site field side {n} unsupported (max {cfg.n})")
        if np.any(self.slices < 1) or np.any(self.slices > cfg.slices):
            raise ValueError("slice counts must lie in [1, cfg.slices]")
        if np.any(self.centres < 0) or np.any(self.centres > cfg.fov):
            raise ValueError("site centres must lie in the field of view")
        vol = torch.as_tensor(index_volume, dtype=torch.float32, device=self.device)
        exp = (cfg.padded, cfg.padded, cfg.slices * cfg.voxels_per_slice)
        if tuple(vol.shape) != exp:
            raise ValueError(f"index volume must have shape {exp}")
        self.volume = vol.contiguous()
        self.windows = site_windows(cfg, self.centres, self.sizes)
        self.bp = dict(beam_parameters(cfg), wvl_um=cfg.wavelength_um)
        self.zern = zernike_weights(cfg)
        d = self.device
        self.site_x = torch.as_tensor(self.windows[:, 0].astype(np.int32), device=d)
        self.site_y = torch.as_tensor(self.windows[:, 1].astype(np.int32), device=d)
        self.site_n = torch.as_tensor(self.sizes.astype(np.int32), device=d)
        self.site_s = torch.as_tensor(self.slices.astype(np.int32), device=d)
        moff = np.concatenate(([0], np.cumsum(self.sizes ** 2)[:-1]))
        self.mask_off = torch.as_tensor(moff.astype(np.int64), device=d)
        self.mask_total = int(np.sum(self.sizes ** 2))
        cost = self.slices * self.sizes ** 2 * np.log2(self.sizes)
        self.orders = {
            "lpt": np.argsort(-cost, kind="stable"),
            "natural": np.arange(P),
            "class_interleaved": np.lexsort((-self.slices, -self.sizes)),
        }
        self.tuned = None
        self.timing = {}
        self._plan = None

    def plan(self, overlap_v4=True):
        """Phase screens, Fresnel tables and base beams.  With overlap_v4, the point-only part of
        the v4 solve (Green matrix, projection, Cholesky: the site centres are all it needs) is
        queued first on a side stream, where its latency-bound kernels run beside this stage's
        small ones; aggregate() picks it up."""
        cfg, d = self.cfg, self.device
        if overlap_v4:
            if getattr(self, "_v4", None) is None:
                self._v4 = V4Ahead(d)
            self._v4.start(self.centres.astype(np.float64))
        lib = nat.lib()
        st = nat.stream_handle(d)
        Z, H = cfg.slices, cfg.padded
        # buffers are allocated once: a captured launch keeps reading the same addresses
        if not hasattr(self, "phase"):
            self.phase = torch.empty((Z, H, H), dtype=torch.float32, device=d)
            self.smax = torch.empty(Z, dtype=torch.float32, device=d)
        nat.check(lib.maomi_phase_prepare(nat.ptr(self.volume), H, H, Z * cfg.voxels_per_slice,
                                          cfg.voxels_per_slice, float(cfg.n0),
                                          float(self.bp["k0dz"]), nat.ptr(self.phase), None, None,
                                          nat.ptr(self.smax), st), "phase prepare")
        check_screen_range(self.smax)
        uniq = sorted({int(n) for n in self.sizes})
        tables = SiteTables(uniq, self.bp["dx"], self.bp["dz"],
                            float(self.bp["wvl"]) / float(self.bp["nr"]), d)
        if not hasattr(self, "tables"):
            self.tables = tables
        else:
            for name in ("edge", "q", "tw"):
                getattr(self.tables, name).copy_(getattr(tables, name))
        # one base beam per (N, S): focus on the last plane, S dz below the entry plane.  The
        # pupil depends on N only; the beams of one N differ in their entry distance, so each
        # N is one pupil and one batched Fresnel step over its depths.
        keys = sorted({(int(n), int(s)) for n, s in zip(self.sizes, self.slices)})
        offs, pos = {}, 0
        for key in keys:
            offs[key] = pos
            pos += key[0] * key[0]
        if not hasattr(self, "bases"):
            self.bases = torch.empty(pos, dtype=torch.complex64, device=d)
            boff = np.array([offs[(int(n), int(s))] for n, s in zip(self.sizes, self.slices)])
            self.base_off = torch.as_tensor(boff.astype(np.int64), device=d)
            self.masks = torch.empty(self.mask_total, dtype=torch.float32, device=d)
            self.energies = torch.empty(len(self.sizes), dtype=torch.float64, device=d)
        for n in sorted({k[0] for k in keys}):
            depths = [s for m, s in keys if m == n]
            c = _CLASSES[n]
            bps = [dict(beam_parameters(replace(cfg, n=n, slices=s, focus_um=s * cfg.dz_um)),
                        wvl_um=cfg.wavelength_um) for s in depths]
            if any(bp[q] != bps[0][q] for bp in bps for q in ("d1", "rad", "objrad", "k", "fl",
                                                              "dx", "wvl", "nr")):
                raise AssertionError("pupil parameters must not depend on the depth")
            pup, _ = make_pupil(n, bps[0], self.zern, d)
            tw = self.tables.tw[c * _SLOT:c * _SLOT + n]
            B = len(depths)
            out = self.bases[offs[(n, depths[0])]:offs[(n, depths[0])] + B * n * n]
            base_beams(pup, bps[0], [bp["dz_focus"] for bp in bps], self.tables.edge2d[n], tw,
                       out.view(B, n, n), d)
        return self

    def _desc(self):
        """The launch descriptor of this run's buffers (the plan supplies the order, grid, block
        shape, workspace and work counter)."""
        cfg = self.cfg
        desc = nat.SitesDesc()
        desc.bases, desc.base_off = self.bases.data_ptr(), self.base_off.data_ptr()
        desc.phase = self.phase.data_ptr()
        desc.edge, desc.q1d = self.tables.edge.data_ptr(), self.tables.q.data_ptr()
        desc.twiddle = self.tables.tw.data_ptr()
        desc.site_x, desc.site_y = self.site_x.data_ptr(), self.site_y.data_ptr()
        desc.site_n, desc.site_s = self.site_n.data_ptr(), self.site_s.data_ptr()
        desc.mask, desc.mask_off = self.masks.data_ptr(), self.mask_off.data_ptr()
        desc.energy = self.energies.data_ptr()
        desc.P, desc.H, desc.W, desc.Z = len(self.sizes), cfg.padded, cfg.padded, cfg.slices
        desc.uniform_n = 0
        for i in range(5):
            desc.class_edge[i] = desc.class_q[i] = desc.class_tw[i] = self.tables.offsets[i]
        return desc

    def _plan_handle(self):
        """The C-ABI plan (maomi_sites_plan_create) over this run's buffers, made once."""
        if getattr(self, "_plan", None) is None:
            h = ctypes.c_void_p()
            nat.check(nat.lib().maomi_sites_plan_create(self._desc(), ctypes.byref(h)),
                      "sites plan")
            self._plan = h
        return self._plan

    def __del__(self):
        h = getattr(self, "_plan", None)
        if h is not None and nat._LIB is not None:
            try:
                nat.lib().maomi_sites_plan_destroy(h)
            except Exception:  # interpreter shutdown: the runtime may already be gone
                pass

    def autotune(self, grids=None, threads=(512, 256), orders=("lpt", "class_interleaved"),
                 schedules=("queue", "static"), reps=2):
        """Time captured candidates (grid, block shape, ordering, schedule) in the plan and
        keep the fastest (maomi_sites_plan_tune)."""
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        grids = [int(g) for g in (grids or (cus, 2 * cus))]
        threads, orders = [int(t) for t in threads], [nat.ORDERS[o] for o in orders]
        schedules = [nat.SCHEDULES[s] for s in schedules]
        arr = lambda v: (ctypes.c_int64 * len(v))(*v)
        n = len(grids) * len(threads) * len(orders) * len(schedules)
        results = (nat.SitesChoice * n)()
        best = nat.SitesChoice()
        nat.check(nat.lib().maomi_sites_plan_tune(
            self._plan_handle(), arr(grids), len(grids), arr(threads), len(threads), arr(orders),
            len(orders), arr(schedules), len(schedules), int(reps),
            nat.stream_handle(self.device), results, ctypes.byref(best)), "sites plan tune")
        name = {v: k for k, v in nat.ORDERS.items()}
        sname = {v: k for k, v in nat.SCHEDULES.items()}
        as_dict = lambda c: {"grid": int(c.grid), "threads": int(c.block_threads),
                             "order": name[int(c.order)], "schedule": sname[int(c.schedule)],
                             "ms": float(c.ms)}
        self.tune_log = [as_dict(c) for c in results]
        self.tuned = as_dict(best)
        return self.tuned

    def propagate(self):
        """Replay the plan's captured launch; without autotune, capture the default choice
        (one block per CU, 512 threads, longest first from the work counter) first."""
        h = self._plan_handle()
        if self.tuned is None:
            cus = torch.cuda.get_device_properties(self.device).multi_processor_count
            nat.check(nat.lib().maomi_sites_plan_capture(
                h, min(cus, len(self.sizes)), 512, nat.ORDERS["lpt"], nat.SCHEDULES["queue"]),
                "sites plan capture")
            self.tuned = {"grid": min(cus, len(self.sizes)), "threads": 512, "order": "lpt",
                          "schedule": "queue", "ms": -1.0}
        nat.check(nat.lib().maomi_sites_plan_replay(h, nat.stream_handle(self.device)),
                  "sites plan replay")
        return self.masks, self.energies

    def site_mask(self, i):
        off, n = int(self.mask_off[i]), int(self.sizes[i])
        return self.masks[off:off + n * n].reshape(n, n)

    def aggregate(self):
        cfg = self.cfg
        prep = self._v4.take() if getattr(self, "_v4", None) is not None else None
        m = v4_masks(self.centres.astype(np.float64), self.energies, (cfg.fov, cfg.fov),
                     self.device, prepared=prep)[..., 0]
        self.mask_raw, self.mask_mean = m, m.mean()
        self.mask = m / self.mask_mean
        return self.mask

    def run(self):
        t = time.perf_counter()
        self.plan()
        torch.cuda.synchronize(self.device)
        self.timing["plan"] = time.perf_counter() - t
        t = time.perf_counter()
        self.propagate()
        torch.cuda.synchronize(self.device)
        self.timing["propagation"] = time.perf_counter() - t
        t = time.perf_counter()
        self.aggregate()
        torch.cuda.synchronize(self.device)
        self.timing["aggregation"] = time.perf_counter() - t
        self.timing["total"] = sum(self.timing[k] for k in ("plan", "propagation", "aggregation"))
        return self
