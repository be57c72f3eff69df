"""GPU parity of the plain optical stage against the CPU oracle (complex64 and FP64).

Bar: rel L2 <= 1e-5 for the FP32 path against the complex64 restatement of
the reference CPU path; the mixed-precision path (FP16 screens) is reported
against the FP64 oracle with the bound written below.
"""

from dataclasses import replace

import numpy as np
import pytest

import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(cfg, sites=None, precision="fp32", block_threads=None):
    import torch
    from maomi_amd.plain import PlainStage
    from maomi_amd.workloads import index_volume, site_offsets, zernike_weights
    vol = index_volume(cfg)
    offs = site_offsets(cfg) if sites is None else sites
    print(f"[plain] N={cfg.n} S={cfg.slices} fov={cfg.fov} sites={len(offs)} {precision} "
          f"threads={block_threads or 'default'}", flush=True)
    st = PlainStage(cfg, vol, sites=offs, precision=precision).plan()
    if block_threads:
        st.block_threads = block_threads
    st.propagate()
    torch.cuda.synchronize()
    return st, vol, offs, zernike_weights(cfg)


def test_config1_fp32_matches_oracle(gpu_lib):
    from maomi_amd.workloads import CONFIG1
    st, vol, offs, zw = _run(CONFIG1)
    st.aggregate()
    ref = O.plain_stage(CONFIG1, vol, offs, zw)
    assert rel_l2(st.pupil.cpu().numpy(), ref["pupil"]) < 1e-6
    assert rel_l2(st.base.cpu().numpy(), ref["base"]) < 2e-6
    assert rel_l2(st.masks.cpu().numpy(), ref["masks"]) <= 1e-5
    assert rel_l2(st.energies.cpu().numpy(), ref["energies"]) <= 1e-5
    assert rel_l2(st.mask_raw.cpu().numpy(), ref["mask_raw"]) <= 1e-5
    np.testing.assert_allclose(float(st.mask.mean()), 1.0, rtol=1e-12)


@pytest.mark.parametrize("threads", [512, 256])
@pytest.mark.parametrize("n,slices,fov,nsites", [(64, 6, 64, 4), (256, 8, 64, 4), (512, 4, 32, 1)])
def test_field_sizes_match_oracle(gpu_lib, n, slices, fov, nsites, threads):
    from maomi_amd.workloads import PlainConfig, site_offsets
    cfg = PlainConfig(fov=fov, n=n, slices=slices, sites_per_axis=2, corr_vox=2.0,
                      zernike_std_rad=0.05)
    st, vol, offs, zw = _run(cfg, site_offsets(cfg)[:nsites], block_threads=threads)
    assert st.block_threads == threads
    ref = O.plain_stage(cfg, vol, offs, zw, with_masks=False)
    assert rel_l2(st.masks.cpu().numpy(), ref["masks"]) <= 1e-5


def test_config2_sampled_sites_fp32(gpu_lib):
    """Full-shape workload (256^2 field, 60 slices, Zernike pupil) on 6 sampled sites."""
    from maomi_amd.workloads import CONFIG2, site_offsets
    rng = np.random.default_rng(CONFIG2.seed + 3)
    sites = site_offsets(CONFIG2)[np.sort(rng.choice(CONFIG2.num_sites, 6, replace=False))]
    st, vol, offs, zw = _run(CONFIG2, sites)
    ref = O.plain_stage(CONFIG2, vol, offs, zw, with_masks=False)
    err = rel_l2(st.masks.cpu().numpy(), ref["masks"])
    print(f"config2 sampled rel L2 vs complex64 oracle: {err:.3e}")
    assert err <= 1e-5


def test_mixed_precision_against_fp64_oracle(gpu_lib):
    from maomi_amd.workloads import CONFIG1
    st, vol, offs, zw = _run(CONFIG1, precision="mixed")
    ref64 = O.plain_stage(CONFIG1, vol.astype(np.float64), offs, zw, dtype=np.complex128,
                          with_masks=False)
    err = rel_l2(st.masks.cpu().numpy(), ref64["masks"])
    print(f"mixed precision rel L2 vs FP64 oracle: {err:.3e}")
    # FP16 screens carry |phi| <= max|phi_z| * 2^-12 per sample; after 16 slices and the
    # fourth power this stays below 2e-3.
    assert err < 2e-3


@pytest.mark.parametrize("X,Y,Zv,vps", [(37, 29, 12, 3), (64, 70, 60, 1), (9, 11, 300, 2)])
def test_phase_prepare_matches_numpy(gpu_lib, X, Y, Zv, vps):
    """Screens phi[z] = float32(k0dz * sum_v (n - n0)) with the sum in FP64, and each slice's
    max |phi|; the shapes cover several voxels per slab, more columns than one block stages
    and ragged last blocks."""
    import torch
    from maomi_amd import _native as nat
    rng = np.random.default_rng(X + Zv)
    vol = (1.35 + 0.02 * rng.standard_normal((X, Y, Zv))).astype(np.float32)
    n0, k0dz = 1.35, 2 * np.pi / 0.92 * 2.0
    Z = Zv // vps
    want = np.float32(k0dz * (vol.astype(np.float64) - n0).reshape(X, Y, Z, vps).sum(-1))
    want = np.moveaxis(want, -1, 0)
    d = torch.as_tensor(vol, device="cuda")
    ph = torch.empty((Z, X, Y), dtype=torch.float32, device="cuda")
    smax = torch.empty(Z, dtype=torch.float32, device="cuda")
    nat.check(nat.lib().maomi_phase_prepare(nat.ptr(d), X, Y, Zv, vps, n0, k0dz, nat.ptr(ph),
                                            None, None, nat.ptr(smax), nat.stream_handle(d.device)),
              "phase prepare")
    got = ph.cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-7)
    np.testing.assert_array_equal(smax.cpu().numpy(), np.abs(got).reshape(Z, -1).max(1))


@pytest.mark.parametrize("n,slices,fov", [(128, 16, 64), (256, 8, 64), (512, 4, 32)])
def test_complex_screens_equal_phase_screens(gpu_lib, n, slices, fov):
    """exp(i phase) precomputed once per voxel gives the per-site sin/cos path's masks."""
    import torch
    from maomi_amd.plain import PlainStage
    from maomi_amd.workloads import PlainConfig, index_volume, site_offsets
    cfg = PlainConfig(fov=fov, n=n, slices=slices, sites_per_axis=2, corr_vox=2.0,
                      zernike_std_rad=0.05)
    vol, offs = index_volume(cfg), site_offsets(cfg)
    print(f"[complex screens] N={n} S={slices} sites={len(offs)}", flush=True)
    out = []
    for screens in ("phase", "complex"):
        st = PlainStage(cfg, vol, sites=offs, screens=screens).plan()
        assert (st.screen is not None) == (screens == "complex")
        st.propagate()
        torch.cuda.synchronize()
        out.append((st.masks.cpu().numpy(), st.energies.cpu().numpy()))
    assert rel_l2(out[1][0], out[0][0]) < 1e-7
    assert rel_l2(out[1][1], out[0][1]) < 1e-7


def test_last_round_split_with_v4_overlap_equals_one_launch(gpu_lib):
    """Sites that do not fill whole rounds of the grid: the whole rounds, then the last round
    with the point-only v4 work on a side stream beside it.  Masks, energies and the v4 mask
    equal the single launch with the whole v4 solve after it, bit for bit, over repeated runs
    with eager work between them."""
    import torch
    from maomi_amd.numerics import V4Ahead
    from maomi_amd.workloads import CONFIG1
    cfg = replace(CONFIG1, sites_per_axis=11)  # 121 sites
    st, vol, offs, zw = _run(cfg)              # grid min(256, 121): one launch
    st.aggregate()
    ref_masks, ref_e, ref = st.masks.clone(), st.energies.clone(), st.mask_raw.clone()
    st.grid = 32                               # 121 = 3 x 32 + 25
    calls = []
    st._v4 = V4Ahead(st.device)
    pts = st.offsets.astype(np.float64)
    st.propagate(tail_hook=lambda: (calls.append(1), st._v4.start(pts)))
    st.aggregate()
    assert calls == [1]
    assert torch.equal(st.masks, ref_masks) and torch.equal(st.energies, ref_e)
    assert torch.equal(st.mask_raw, ref)
    for _ in range(3):
        st.run()
        assert torch.equal(st.masks, ref_masks) and torch.equal(st.energies, ref_e)
        assert torch.equal(st.mask_raw, ref)
        st.energies.sum().item()


def _config2_sample(k):
    from maomi_amd.workloads import CONFIG2, site_offsets
    rng = np.random.default_rng(CONFIG2.seed + 3)
    return site_offsets(CONFIG2)[np.sort(rng.choice(CONFIG2.num_sites, k, replace=False))]


def test_config4_full_shape_against_fp64_oracle(gpu_lib):
    """Mixed precision at the 4225-site workload's shape (256^2 field, 60 slices, Zernike
    pupil): FP16 screens with a per-slice max-abs scale, FP32 fields and FFTs, against the FP64
    oracle on 8 seeded sampled sites.  Bar from the measured 1.2e-4 (64 sites, worst 1.9e-4):
    2e-4 over the sample, 3e-4 per site."""
    from maomi_amd.workloads import CONFIG2
    st, vol, offs, zw = _run(CONFIG2, _config2_sample(8), precision="mixed")
    assert st.phase is None and st.phase16 is not None
    ref = O.plain_stage(CONFIG2, vol.astype(np.float64), offs, zw, dtype=np.complex128,
                        with_masks=False)["masks"]
    got = st.masks.cpu().numpy()
    per = [rel_l2(g, r) for g, r in zip(got, ref)]
    err = rel_l2(got, ref)
    print(f"config4 rel L2 vs FP64 oracle: all {err:.3e}, worst site {max(per):.3e}", flush=True)
    assert err <= 2e-4 and max(per) <= 3e-4


def test_config2_fp32_against_fp64_oracle(gpu_lib):
    """The FP32 path at the 4225-site workload's shape against the FP64 oracle.  The float32
    pupil (a focusing chirp of ~5e4 rad, computed in the reference's float32 order) puts the
    complex64 reference path itself ~2.2e-5 from FP64, and this path follows it (bar 4e-5).
    Given the FP64 oracle's entry beam, the FP32 propagation over 60 slices is within 1.5e-5
    of the FP64 oracle (the complex64 CPU propagation measures 9.8e-6 there)."""
    import torch
    from maomi_amd.workloads import CONFIG2
    st, vol, offs, zw = _run(CONFIG2, _config2_sample(4))
    ref = O.plain_stage(CONFIG2, vol.astype(np.float64), offs, zw, dtype=np.complex128,
                        with_masks=False)
    err = rel_l2(st.masks.cpu().numpy(), ref["masks"])
    st.base.copy_(torch.as_tensor(ref["base"].astype(np.complex64), device=st.device))
    st.propagate()
    torch.cuda.synchronize()
    err64 = rel_l2(st.masks.cpu().numpy(), ref["masks"])
    print(f"config2 FP32 vs FP64 oracle: {err:.3e}; from the FP64 beam: {err64:.3e}", flush=True)
    assert err < 4e-5
    assert err64 < 1.5e-5


def test_fp16_screens_match_oracle_on_the_same_screens(gpu_lib):
    """The FP16-screen kernels against the complex64 oracle given the same dequantized screens
    (phi = float(phase16) * slice scale, as the kernel reads them), at the FP32 bar.  With index
    std 0.035 about one phase sample in a thousand lies beyond pi/2, so a wave's 16 x 64 phases
    of a job are all below it about half the time: part of the jobs take the full sin/cos range
    reduction and part skip it, and both paths of the kernel are compared."""
    import torch
    from maomi_amd.workloads import CONFIG1, site_offsets, zernike_weights, index_volume
    cfg = replace(CONFIG1, dn_std=0.035)
    st, vol, offs, zw = _run(cfg, precision="mixed")
    ph16 = st.phase16.float().cpu().numpy()
    scale = st.scale.cpu().numpy()
    phase = (ph16 * scale[:, None, None]).astype(np.float32)
    big = np.mean(np.abs(phase) > 1.5707)
    print(f"fp16 screens: max |phi| {np.abs(phase).max():.2f} rad, {big:.1%} of samples "
          f"beyond pi/2", flush=True)
    assert 1e-4 < big < 0.02
    base, _ = O.plain_beam(cfg, zw)
    ref = O.plain_stage_masks(base, phase, np.asarray(offs, int),
                              wvl=np.float32(cfg.wavelength_um * 1e-6), nr=np.float32(cfg.n0),
                              dx=np.float32(cfg.dx_um * 1e-6), dz=cfg.dz_um * 1e-6)
    err = rel_l2(st.masks.cpu().numpy(), np.stack(ref))
    print(f"fp16 screens vs oracle on the same screens: {err:.3e}", flush=True)
    assert err <= 1e-5
