"""Pin the oracle's plain optical stage (plain_stage_masks, plain_phase, plain_beam, plain_stage).

These functions define the generic split-step workload that every GPU test of the
plain, uneven and mixed-precision launches is judged against, so each piece is
checked here against something independent of its own code:
  * the energy budget of a unitary propagation (order 1, no edge window);
  * free propagation in closed form (zero screens: every saved plane is one exact
    transfer-function step of k dz, computed here in FP64 without the oracle);
  * covariance under a joint shift of the volume and the site offsets (an x/y
    swap or an off-by-one in the window fails it);
  * the golden-pinned segment A of the oracle's propagate at cut = 0 (per-plane
    profiles, tail weights and per-pixel tail images, bit for bit);
  * the reference's own pupil vectors for plain_beam's pupil grid and pitch, and
    the focus of its beam at the stated depth;
  * a direct voxel sum for plain_phase.
"""

from dataclasses import replace

import numpy as np
import pytest

import oracle as O
from tests.conftest import rel_l2
from tests.golden.make_golden import golden_raw_fields

WVL, NR, DX = np.float32(0.92e-6), np.float32(1.35), np.float32(0.5e-6)


def _random_case(N=32, S=5, H=48, seed=3, amp=0.4):
    rng = np.random.default_rng(seed)
    base = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))).astype(np.complex64)
    base /= np.float32(np.linalg.norm(base))
    phase = (amp * rng.standard_normal((S, H, H))).astype(np.float32)
    return base, phase


def test_order1_without_edge_conserves_energy_per_plane():
    """|Q| = 1 and |exp(i phi)| = 1 make each step unitary, so every saved plane carries
    |base|^2 and the trapezoid over S + 1 planes (weights 1/2, 1, ..., 1, 1/2) sums to S."""
    base, phase = _random_case()
    S = phase.shape[0]
    offs = np.array([[0, 0], [16, 3], [5, 11]])
    e0 = float(np.sum(np.abs(base.astype(np.complex128)) ** 2))
    for dtype, tol in ((np.complex64, 2e-6), (np.complex128, 1e-12)):
        masks = O.plain_stage_masks(base.astype(dtype), phase, offs, wvl=WVL, nr=NR, dx=DX,
                                    dz=2e-6, edge=False, dtype=dtype, order=1)
        for m in masks:
            assert abs(m.sum(dtype=np.float64) - S * e0) < tol * S * e0
    # with the edge window the field loses energy at the border
    masks = O.plain_stage_masks(base, phase, offs, wvl=WVL, nr=NR, dx=DX, dz=2e-6, edge=True,
                                order=1)
    assert all(m.sum() < S * e0 for m in masks)


def _free_closed_form(u, dist, dx=float(DX), wvl=float(WVL), nr=float(NR)):
    """Paraxial transfer function exp(-i pi lambda' z (fx^2 + fy^2)) in FP64, applied with
    numpy's FFT in natural frequency order (no shifts): the textbook free-space step."""
    n = u.shape[0]
    f = np.fft.fftfreq(n, d=dx)
    H = np.exp(-1j * np.pi * (wvl / nr) * dist * (f[:, None] ** 2 + f[None, :] ** 2))
    return np.fft.ifft2(np.fft.fft2(u) * H)


def test_zero_screens_give_free_propagation_planes():
    """With zero screens and no edge, saved plane k is the free field at k dz; the mask is the
    trapezoid of |u_k|^4 over k = 0..S."""
    base, _ = _random_case(N=32)
    S, dz = 6, 3e-6
    zero = np.zeros((S, 40, 40), np.float32)
    masks, planes = O.plain_stage_masks(base.astype(np.complex128), zero, [[4, 2]], wvl=WVL,
                                        nr=NR, dx=DX, dz=dz, edge=False, dtype=np.complex128,
                                        return_planes=True)
    assert planes[0].shape == (32, 32, S + 1)
    want = np.zeros((32, 32))
    for k in range(S + 1):
        uk = _free_closed_form(base.astype(np.complex128), k * dz)
        assert rel_l2(planes[0][..., k], uk) < 1e-12, k
        want += (0.5 if k in (0, S) else 1.0) * np.abs(uk) ** 4
    assert rel_l2(masks[0], want) < 1e-12


def test_masks_follow_a_joint_shift_of_volume_and_offsets():
    """Window semantics: site (x, y) reads phase[:, x:x+N, y:y+N].  Shifting the volume by
    (a, b) with a != b and the offsets by the same (a, b) leaves every mask unchanged."""
    base, phase = _random_case(N=16, S=4, H=40)
    offs = np.array([[0, 0], [7, 2], [11, 19]])
    a, b = 3, 5
    shifted = np.zeros((4, 48, 48), np.float32)
    shifted[:, a:a + 40, b:b + 40] = phase
    m0 = O.plain_stage_masks(base, phase, offs, wvl=WVL, nr=NR, dx=DX, dz=2e-6)
    m1 = O.plain_stage_masks(base, shifted, offs + [a, b], wvl=WVL, nr=NR, dx=DX, dz=2e-6)
    for p, q in zip(m0, m1):
        np.testing.assert_array_equal(p, q)
    # the swapped shift reads other voxels: masks change
    m2 = O.plain_stage_masks(base, shifted, offs + [b, a], wvl=WVL, nr=NR, dx=DX, dz=2e-6)
    assert min(rel_l2(p, q) for p, q in zip(m2, m0)) > 1e-3


def test_masks_equal_propagate_segment_a_at_cut_zero():
    """On a geometry with n2 = n (no crop), an integral number of A slabs and cut = 0, the
    oracle's propagate (golden-pinned against the reference) runs exactly the plain workload in
    segment A: fresnel over the site's window of exp(i phase) * edge with constant pitch d2 and
    spacing prop_um, all planes saved.  Its top tail then gives, per site, the per-plane
    profiles ss^2 sum_xy |u_p|^4, the weight prop_um ss^2 sum(mask) and the tail image
    ss^2 resize(translate(ss^2 mask)) -- all equal to the plain masks' values."""
    cfg = O.OracleConfig(volume_um=(8, 8, 10), depth_um=5, psf_um=(4, 4, 4), tail_um=40,
                         vres=1, propcrop=False, prop_um=2)
    shapes = O.field_shapes(cfg)
    g = O.geometry(cfg, [s[-1] for s in shapes])
    assert g.n == g.n2 and g.cut == 0 and np.allclose(np.diff(g.z[0]), cfg.prop_um * 1e-6)
    raw = golden_raw_fields(shapes, seed=7)
    pts = np.array([[0.0, 0.0], [8.0, 3.5], [2.5, 6.0]])
    res = O.propagate(cfg, raw, pts)
    r, ss, rdt = cfg.vres, cfg.ss, np.float32
    arg = O.resize_xy(np.asarray(raw[0], rdt), ss) * rdt(cfg.n_diff * 2 * np.pi /
                                                         (cfg.wavelength_um * r))
    sg = O.edge_apodization(g.n)
    dz0 = np.float32(g.fl - (cfg.depth_um + cfg.volume_um[2] / 2) * 1e-6)
    base = O.base_field(O.pupil(cfg), g.wvl, g.d1, g.d2, dz0, g.nr, sg, norm="exact")
    offs = np.round(pts * r * ss).astype(int)
    masks, planes = O.plain_stage_masks(base, np.moveaxis(arg, -1, 0), offs, wvl=g.wvl, nr=g.nr,
                                        dx=g.d2, dz=cfg.prop_um * 1e-6, return_planes=True)
    top = res["psfT"]
    for i in range(len(pts)):
        prof = rdt(ss ** 2) * (np.abs(planes[i]) ** 4).sum((0, 1))
        np.testing.assert_allclose(prof, top["profiles_per_point"][i], rtol=1e-6)
        w = float((rdt(ss ** 2) * masks[i]).sum()) * cfg.prop_um
        assert abs(w - top["weights_per_point"][i]) < 1e-6 * w
        img = rdt(ss ** 2) * O.resize_xy(O.translate_xy(rdt(ss ** 2) * masks[i], ss / 2 - 0.5),
                                         1 / ss)
        np.testing.assert_allclose(img, top["convmask_per_point"][i], rtol=1e-6,
                                   atol=1e-6 * np.abs(img).max())


def test_plain_beam_pupil_is_the_reference_pupil(golden):
    """plain_beam's pupil on an N-sample grid with pitch max(beam size)/N is generate_ba's, so at
    the scene geometry's N = 92 it reproduces the reference's own Zernike pupil vectors."""
    from maomi_amd.workloads import PlainConfig
    w = tuple(golden["pupil_scene_zernike_w"])
    cfg = PlainConfig(n=92, NA=0.6, objNA=0.8, n0=1.35, wavelength_um=0.92, focal_length_mm=4.5)
    _, u = O.plain_beam(cfg, w)
    assert rel_l2(u, golden["pupil_scene_zernike"]) < 1e-5
    # without the Zernike weights the pupils differ: the weights reach the beam
    _, u0 = O.plain_beam(cfg, ())
    assert rel_l2(u0, golden["pupil_scene_zernike"]) > 1e-2


def test_plain_beam_focuses_at_the_stated_depth():
    """The entry field is the pupil carried over fl - focus, so free propagation over the
    slices brings the beam to its focus at depth `focus`: the peak of |u|^2 over the saved
    planes lies at plane focus / dz, and the beam has unit L2 norm."""
    from maomi_amd.workloads import PlainConfig
    for focus_um, S in ((10.0, 12), (6.0, 12)):
        cfg = PlainConfig(n=128, slices=S, focus_um=focus_um)
        base, _ = O.plain_beam(cfg, (), np.complex128)
        assert abs(np.linalg.norm(base) - 1) < 1e-12
        peaks = [np.abs(_free_closed_form(base, k * cfg.dz_um * 1e-6)).max() for k in range(S + 1)]
        assert int(np.argmax(peaks)) == round(focus_um / cfg.dz_um)


def test_plain_phase_is_the_slab_sum():
    """phi[z, x, y] = k0 dz sum over the slab's voxels of (n - n0), slices first."""
    rng = np.random.default_rng(4)
    X, Y, vps, Z = 5, 4, 3, 2
    vol = (1.35 + 0.02 * rng.standard_normal((X, Y, Z * vps))).astype(np.float32)
    k0dz = 2 * np.pi / 0.92 * 2.0 / vps
    got = O.plain_phase(vol, 1.35, k0dz, vps, np.complex128)
    assert got.shape == (Z, X, Y)
    for z in range(Z):
        for x in range(X):
            for y in range(Y):
                s = sum(float(vol[x, y, z * vps + v]) - 1.35 for v in range(vps))
                assert abs(got[z, x, y] - k0dz * s) < 1e-12


def test_plain_stage_mask_is_v4_of_site_energies():
    """plain_stage's FOV mask: v4 of the per-site energies at the offsets, normalised to mean
    1; with one site the mask is constant (the reference's single-site contract)."""
    from maomi_amd.workloads import PlainConfig, index_volume, site_offsets
    cfg = PlainConfig(fov=16, n=64, slices=3, sites_per_axis=2)
    vol = index_volume(cfg)
    offs = site_offsets(cfg)
    res = O.plain_stage(cfg, vol, offs)
    np.testing.assert_allclose(res["energies"], [m.sum(dtype=np.float64) for m in res["masks"]])
    # v4 reproduces its nodes: the pixel centre nearest each site holds about its energy
    assert abs(res["mask"].mean() - 1) < 1e-12
    one = O.plain_stage(cfg, vol, offs[:1])
    assert np.ptp(one["masks"][0]) > 0
    sub = O.plain_stage(cfg, vol, offs, sites=[2])
    np.testing.assert_array_equal(sub["masks"][0], res["masks"][2])


@pytest.mark.slow
def test_config2_float32_error_against_fp64_is_the_pupil_chirp():
    """At the 4225-site workload's shape (256^2 field, 60 slices, Zernike pupil), the complex64
    restatement of the reference CPU path is about 2.2e-5 from the FP64 oracle.  Nearly all of
    it is the float32 pupil (its focusing chirp reaches ~5e4 rad): the same float32
    propagation started from the FP64 beam is within 1e-5."""
    from maomi_amd.workloads import CONFIG2, index_volume, site_offsets, zernike_weights
    cfg = CONFIG2
    vol = index_volume(cfg)
    zw = zernike_weights(cfg)
    rng = np.random.default_rng(cfg.seed + 3)
    offs = site_offsets(cfg)[np.sort(rng.choice(cfg.num_sites, 2, replace=False))]
    a = O.plain_stage(cfg, vol, offs, zw, with_masks=False)
    b = O.plain_stage(cfg, vol.astype(np.float64), offs, zw, dtype=np.complex128,
                      with_masks=False)
    err32 = rel_l2(a["masks"], b["masks"])
    assert 1e-5 < err32 < 4e-5
    assert rel_l2(a["pupil"], b["pupil"]) > 1e-4
    k0dz = 2 * np.pi / cfg.wavelength_um * cfg.dz_um
    ph = O.plain_phase(vol, cfg.n0, k0dz, 1, np.complex64)
    m = O.plain_stage_masks(b["base"].astype(np.complex64), ph, offs, wvl=WVL, nr=NR, dx=DX,
                            dz=cfg.dz_um * 1e-6)
    assert rel_l2(np.stack(m), b["masks"]) < 1e-5
