"""Synthetic inputs of the plain optical stage (input generation, outside the timed stage).

A plain-stage run is a refractive-index volume n(x, y, z) on a voxel grid,
padded by one field side N around the field of view, a grid of scan sites,
and a focused beam.  Each site propagates the beam through S slices of the
volume by split-step Fresnel propagation (the constant-pitch case of
maomi/numerics.py:156-210) and accumulates its two-photon map sum |E|^4.

Phase per slice and voxel: phi = k0 * dz * (n - n0) (maomi/optics.py:232,
phase_scale * raw with raw = (n - n0)/n_diff summed over the slab).
"""

from dataclasses import dataclass, field, replace

import numpy as np


@dataclass(frozen=True)
class PlainConfig:
    fov: int = 128                 # field of view side in voxels (X = Y)
    n: int = 128                   # per-site field side N
    slices: int = 16               # screens S (slices of the volume)
    sites_per_axis: int = 3
    dx_um: float = 0.5             # lateral voxel / sample pitch
    dz_um: float = 2.0             # axial voxel pitch = slice thickness
    wavelength_um: float = 0.92
    n0: float = 1.35
    NA: float = 0.6                # beam 1/e radius aperture
    objNA: float = 0.8             # circular pupil aperture
    focal_length_mm: float = 4.5
    focus_um: float = -1.0         # focal depth below the entry plane; < 0 -> S*dz/2
    corr_vox: float = 2.0          # Gaussian correlation length of the index field (voxels)
    dn_std: float = 0.02
    seed: int = 20260927
    zernike_std_rad: float = 0.0   # Noll modes 4..15, c_j ~ N(0, std) radians
    voxels_per_slice: int = 1

    @property
    def focus(self):
        return self.focus_um if self.focus_um >= 0 else self.slices * self.dz_um / 2

    @property
    def padded(self):
        return self.fov + self.n

    @property
    def num_sites(self):
        return self.sites_per_axis ** 2

    def bytes_per_site(self, half=False):
        """Algorithmic HBM bytes of one site: S (8N^2 read + 8N^2 write + phase) + 8N^2 + 4N^2."""
        n2 = self.n * self.n
        ph = 2 if half else 4
        return self.slices * (16 + ph) * n2 + 12 * n2 + (4 * self.slices if half else 0)

    def flops_per_site(self):
        n2 = self.n * self.n
        return self.slices * (2 * 5 * n2 * np.log2(n2) + 20 * n2)


# The plain-stage workloads: 128^2 x 16 on 3x3 sites, and 512^2 x 60 on 65x65 sites
CONFIG1 = PlainConfig()
CONFIG2 = PlainConfig(fov=512, n=256, slices=60, sites_per_axis=65, corr_vox=3.0,
                      zernike_std_rad=0.05)


def zernike_weights(cfg):
    """Noll j = 4..15 aberration, c_j ~ N(0, std) rad, entered as c_j / (2 pi n0) waves."""
    if cfg.zernike_std_rad <= 0:
        return ()
    rng = np.random.default_rng(cfg.seed + 1)
    c = rng.normal(0.0, cfg.zernike_std_rad, 12)
    w = np.zeros(15)
    w[3:15] = c / (2 * np.pi * cfg.n0)
    return tuple(float(v) for v in w)


def index_volume(cfg, dtype=np.float32):
    """n0 + Gaussian random field (std dn_std, Gaussian correlation corr_vox voxels), seeded.

    Shape (fov + N, fov + N, slices * voxels_per_slice), axes X, Y, Z.
    """
    shape = (cfg.padded, cfg.padded, cfg.slices * cfg.voxels_per_slice)
    rng = np.random.default_rng(cfg.seed)
    white = rng.standard_normal(shape, dtype=np.float32)
    spec = np.fft.rfftn(white)
    axes = []
    for d, size in enumerate(shape):
        f = np.fft.fftfreq(size) if d < 2 else np.fft.rfftfreq(size)
        axes.append(np.exp(-2 * (np.pi * cfg.corr_vox * f) ** 2))
    spec *= axes[0][:, None, None] * axes[1][None, :, None] * axes[2][None, None, :]
    field_ = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2))
    field_ *= cfg.dn_std / field_.std()
    return (cfg.n0 + field_).astype(dtype)


def site_offsets(cfg):
    """Window top-left corners (= site centres in FOV pixels) on a uniform grid incl. edges."""
    ticks = np.round(np.linspace(0, cfg.fov, cfg.sites_per_axis)).astype(np.int64)
    xx, yy = np.meshgrid(ticks, ticks, indexing="ij")
    return np.stack((xx.ravel(), yy.ravel()), -1)


def beam_parameters(cfg):
    """generate_ba's pupil quantities for an explicit N-sample grid (_optical_math.py:232-256)."""
    fl = float(cfg.focal_length_mm / 1000.0)
    r = np.ceil(np.tan(np.arcsin(cfg.objNA / cfg.n0)) * fl * 1e6 * 1.5) * 2.0
    beam = np.array([r, r, 0], dtype=np.float32)
    d1 = float(max(beam) / 1e6 / np.float32(cfg.n))
    return dict(
        fl=fl, d1=d1,
        rad=float(np.tan(np.arcsin(cfg.NA / cfg.n0)) * fl),
        objrad=float(np.tan(np.arcsin(cfg.objNA / cfg.n0)) * fl),
        k=2 * cfg.n0 * np.pi / float(cfg.wavelength_um * 1e-6),
        wvl=np.float32(cfg.wavelength_um * 1e-6), nr=np.float32(cfg.n0),
        dx=np.float32(cfg.dx_um * 1e-6), dz=cfg.dz_um * 1e-6,
        dz_focus=np.float32(fl - cfg.focus * 1e-6),
        k0dz=2 * np.pi / cfg.wavelength_um * cfg.dz_um / cfg.voxels_per_slice,
    )
