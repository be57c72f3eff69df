"""The optical stage as a PSF provider (reference protocol: maomi/rana/providers.py:10-20).

``OpticalStageProvider(config, volume, ...).build_field()`` runs the GPU optical stage
(``maomi_amd.fixed.run_fixed``) on a site lattice and returns the per-site local PSFs as a
``FullFieldPSF``, the way the reference's native attachment turns ``psf_3d`` into a field
(maomi/rana/adapters/naomi.py:38-55): the lattice must be complete and rectangular, kernels
are |psf_3d| ordered by (x, y), and every kernel is scaled to the run's mean-PSF energy
(the local gain stays in the source mask).  Any object with ``build_field()`` is accepted by
``maomi_amd.fields.resolve_field``; no registration.
"""

import numpy as np

from .fields import FullFieldPSF


def native_field(optics, *, config=None, voxel_size_um=None, wavelength_nm=None):
    """FullFieldPSF of a run's per-site PSFs (``points_um``, ``psf_3d``, ``psf``)."""
    if voxel_size_um is None:
        if config is None:
            raise ValueError("Specify voxel_size_um when building a field without OpticalConfig")
        voxel_size_um = (1.0 / config.vres,) * 3
    if wavelength_nm is None:
        wavelength_nm = config.wavelength_um * 1000 if config is not None else 920.0
    pts = np.asarray(optics["points_um"], float)
    xs, ys = np.unique(pts[:, 0]), np.unique(pts[:, 1])
    if len(pts) != len(xs) * len(ys) or len(np.unique(pts, axis=0)) != len(pts):
        raise ValueError("Native RANA requires a complete rectangular propagation grid")
    k = np.abs(np.asarray(optics["psf_3d"], float))[np.lexsort((pts[:, 1], pts[:, 0]))]
    energy = k.sum(axis=(1, 2, 3))
    if (energy <= 0).any():
        raise ValueError("Every native local PSF must have positive energy")
    k *= float(np.asarray(optics["psf"]).sum()) / energy[:, None, None, None]
    return FullFieldPSF(k.reshape(len(xs), len(ys), *k.shape[1:]), xs, ys, voxel_size_um,
                        wavelength_nm, "excitation",
                        {"source": "native per-position propagation (MI355X optical stage)",
                         "offset_convention": "MATLAB",
                         "local_gain": "retained in existing source mask"})


class OpticalStageProvider:
    """build_field() -> FullFieldPSF from one GPU optical-stage run (no caching: a provider
    is resolved once, at preparation)."""

    def __init__(self, config, volume, *, count_per_axis=65, device="cuda", **run_kwargs):
        self.config, self.volume = config, volume
        self.count_per_axis, self.device = int(count_per_axis), device
        self.run_kwargs = dict(run_kwargs)
        self.result = None

    def build_field(self):
        from .fixed import run_fixed
        kw = dict(self.run_kwargs)
        kw.setdefault("render_image", False)
        self.result = run_fixed(self.config, self.volume, count_per_axis=self.count_per_axis,
                                device=self.device, **kw)
        return native_field(self.result, config=self.config)


class OpticalStagePSF:
    """The GPU optical stage as a 2-D PSF provider (maomi/psf.py:66-71): a callable
    (x_um, y_um) -> SampledPSF.

    The kernel at a site is the stage's focal two-photon PSF (``OpticalPlan.evaluate``'s
    ``psfs``, axes X, Y, Z at 1 / vres um) reduced over Z -- the axial sum by default, or the
    central plane with ``projection="focus"`` -- as magnitudes (|psf|, as the native RANA
    attachment takes them), normalised to unit sum.  Pitch 1 / vres um, wavelength the
    excitation wavelength.  ``prefetch(x_um, y_um)`` evaluates every node of a grid in one
    persistent launch, so ``PSFField.from_provider`` costs one launch; nodes are kept, and a
    node outside every prefetched grid is evaluated on its own.  Sites must land on the
    supersampled lattice (the plan's own check)."""

    def __init__(self, config, raw_fields, *, uin=None, device="cuda", projection="sum"):
        from .batched import OpticalPlan
        if projection not in ("sum", "focus"):
            raise ValueError("projection must be 'sum' or 'focus'")
        self.config, self.projection = config, projection
        self.plan = OpticalPlan(config, raw_fields, uin=uin, device=device)
        self.pixel_size_um = (1.0 / config.vres, 1.0 / config.vres)
        self.wavelength_nm = float(config.wavelength_um) * 1000.0
        self.kernels = {}
        self.launches = 0

    def prefetch(self, x_um, y_um):
        xs, ys = np.asarray(x_um, float).ravel(), np.asarray(y_um, float).ravel()
        pts = np.stack(np.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
        todo = np.array([p for p in pts if (float(p[0]), float(p[1])) not in self.kernels])
        if not len(todo):
            return
        psfs = np.abs(np.asarray(self.plan.evaluate(todo)["psfs"], np.float64))
        self.launches += 1
        k = psfs.sum(-1) if self.projection == "sum" else psfs[..., psfs.shape[-1] // 2]
        for p, kern in zip(todo, k):
            self.kernels[(float(p[0]), float(p[1]))] = kern

    def __call__(self, x_um, y_um):
        from .psf import SampledPSF
        key = (float(x_um), float(y_um))
        if key not in self.kernels:
            self.prefetch([key[0]], [key[1]])
        return SampledPSF.normalized(
            self.kernels[key], self.pixel_size_um, self.wavelength_nm,
            {"source": "MI355X optical stage (two-photon focal PSF)",
             "projection": self.projection, "x_um": key[0], "y_um": key[1]})
