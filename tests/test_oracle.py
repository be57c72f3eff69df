"""Pin the CPU oracle: reference golden vectors, closed forms and brute force."""

import json
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O
from tests.conftest import rel_l2
from tests.golden.make_golden import golden_raw_fields

GOLD = Path(__file__).resolve().parent / "golden"


def _cfg(**kw):
    return O.OracleConfig(**kw)


# ----------------------------------------------------------------- geometry and helpers
def test_geometry_matches_reference_on_all_configs():
    geo = json.loads((GOLD / "geometry.json").read_text())
    assert len(geo) >= 7
    for name, g in geo.items():
        cfg = _cfg(**{k: tuple(v) if isinstance(v, list) else v for k, v in g["config"].items()})
        shapes = O.field_shapes(cfg)
        assert [list(s) for s in shapes] == g["shapes"], name
        if "error" in g:
            with pytest.raises(ValueError, match="Incompatible|Top tail"):
                O.geometry(cfg, [s[-1] for s in shapes])
            continue
        mine = O.geometry(cfg, [s[-1] for s in shapes])
        assert (mine.n, mine.n2, mine.cut) == (g["n"], g["n2"], g["cut"]), name
        assert [float(mine.fl), float(mine.d1), float(mine.d2)] == [g["fl"], g["d1"], g["d2"]]
        assert float(mine.wvl) == g["wvl"] and float(mine.nr) == g["nr"]
        for za, zb in zip(mine.z, g["z"]):
            np.testing.assert_array_equal(za, zb)
        assert list(mine.focal_shape) == g["focal_shape"]
        assert [mine.crop.start, mine.crop.stop] == g["crop"]
        assert [[s.start, s.stop] for s in mine.focal_crop] == g["focal_crop"]


def test_top_tail_without_extent_is_rejected():
    cfg = _cfg(volume_um=(8.0, 8.0, 4.0), depth_um=20.0, vres=1.0, tail_um=1.0, psf_um=(4, 4, 4))
    shapes = O.field_shapes(cfg)
    with pytest.raises(ValueError, match="Top tail has no extent"):
        O.geometry(cfg, [s[-1] for s in shapes])


def test_center_slice_errors():
    assert O.center_slice(10, 4) == slice(3, 7)
    for n, m in ((4, 6), (9, 4)):
        with pytest.raises(ValueError, match="Incompatible centered grids"):
            O.center_slice(n, m)


def test_edge_window(golden):
    for size in (8, 14, 36):
        np.testing.assert_allclose(O.edge_apodization(size), golden[f"edge_{size}"], rtol=2e-7,
                                   atol=1e-30)


def test_lattices_and_axial(golden):
    demo = _cfg(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)
    np.testing.assert_array_equal(O.fixed_locations(demo), golden["fixed_demo_65"])
    scene = _cfg(volume_um=(32, 32, 20), depth_um=10, psf_um=(4, 4, 4), tail_um=5, vres=1)
    np.testing.assert_array_equal(O.fixed_locations(scene, 3), golden["fixed_scene_3"])
    tf = _cfg(scaling="temporal-focusing")
    np.testing.assert_allclose(O.axial_profile(tf, 20), golden["axial_tf_focus"], rtol=1e-7)
    np.testing.assert_allclose(O.axial_profile(tf, 7, "top"), golden["axial_tf_top"], rtol=1e-7)
    np.testing.assert_allclose(O.axial_profile(tf, 5, "bottom"), golden["axial_tf_bottom"],
                               rtol=1e-7)
    for count in (1, 3.5, True):
        with pytest.raises(ValueError):
            O.fixed_locations(demo, count)


# ----------------------------------------------------------------- resampling
def test_resize_translate_match_reference(golden):
    a = golden["resize_in"]
    for s, tag in ((2, "2"), (0.5, "half"), (1 / 3, "third")):
        np.testing.assert_allclose(O.resize_xy(a, s), golden[f"resize_{tag}"], rtol=2e-6,
                                   atol=2e-6)
    b = golden["translate_in"]
    for s, tag in ((0.5, "half"), (1.0, "one"), (1.5, "onehalf")):
        np.testing.assert_allclose(O.translate_xy(b, s), golden[f"translate_{tag}"], atol=2e-6)


def test_resize_reproduces_constants_and_interior_ramps():
    const = np.full((12, 10), 3.25)
    np.testing.assert_allclose(O.resize_xy(const, 2), 3.25, rtol=1e-13)
    np.testing.assert_allclose(O.resize_xy(const, 0.5), 3.25, rtol=1e-13)
    ramp = np.add.outer(np.arange(16.0), np.zeros(16))
    up = O.resize_xy(ramp, 2)
    # interior samples of a linear ramp are exact for the Keys kernel at output centres
    centres = (np.arange(32) + 0.5) / 2 - 0.5
    np.testing.assert_allclose(up[4:-4, 4], centres[4:-4], atol=1e-12)
    # Keys kernel values at known offsets
    np.testing.assert_allclose(O.cubic(np.array([0, 0.5, 1, 1.5, 2])),
                               [1, 0.5625, 0, -0.0625, 0], atol=1e-15)


def test_resample_matrix_equals_translate_then_resize():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((24, 20))
    mx, my = O.resample_matrix(24, 2), O.resample_matrix(20, 2)
    ref = O.resize_xy(O.translate_xy(a, 0.5), 0.5)
    np.testing.assert_allclose(mx @ a @ my.T, ref, atol=1e-12)


# ----------------------------------------------------------------- fresnel
def test_fresnel_matches_reference(golden):
    u, phi = golden["fresnel_u"], golden["fresnel_phi"]
    o, saved = O.fresnel(u, np.float32(0.92e-6), golden["fresnel_dx"], golden["fresnel_z"], phi,
                         np.float32(1.35), return_all=True, crop=(slice(8, 24), slice(4, 28)),
                         save_from=1)
    assert rel_l2(o, golden["fresnel_out"]) < 2e-6
    assert rel_l2(saved, golden["fresnel_saved"]) < 2e-6
    sg = O.edge_apodization(32)
    so = O.fresnel(u, np.float32(0.92e-6), [np.float32(4.4e-5), np.float32(0.25e-6)],
                   [0, np.float32(4.48e-3)], np.stack((sg, sg), -1), np.float32(1.35))
    assert rel_l2(so, golden["fresnel_scaled_out"]) < 5e-5


def _free(u, dz_list, dx=0.5e-6, wvl=0.92e-6, nr=1.35):
    z = np.r_[0, np.cumsum(dz_list)]
    phi = np.ones(u.shape + (len(z) - 1,), u.dtype)
    return O.fresnel(u, wvl, np.full(len(z), dx), z, phi, nr)


def test_fresnel_unitary_and_semigroup():
    rng = np.random.default_rng(11)
    u = (rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32)))
    one = _free(u, [3e-6] * 4)
    assert abs(np.linalg.norm(one) / np.linalg.norm(u) - 1) < 1e-12
    big = _free(u, [12e-6])
    np.testing.assert_allclose(one, big, atol=1e-12 * np.abs(u).max())


def test_fresnel_gaussian_beam_closed_form():
    N, dx, lam, nr = 128, 0.25e-6, 0.92e-6, 1.35
    x = (np.arange(N) - N / 2) * dx
    w0 = 2e-6
    u = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / w0 ** 2).astype(np.complex128)
    z = 10e-6
    out = _free(u, [z], dx=dx, wvl=lam, nr=nr)
    lp = lam / nr
    zr = np.pi * w0 ** 2 / lp
    # transfer function exp(-i pi lambda' z f^2) <-> exp(+ikz) convention: q(z) = z - i zR,
    # u(z) = (q0/qz) exp(+i k r^2 / (2 qz)), global exp(ikz) dropped
    q0, qz = -1j * zr, z - 1j * zr
    k = 2 * np.pi / lp
    r2 = x[:, None] ** 2 + x[None, :] ** 2
    ana = (q0 / qz) * np.exp(1j * k * r2 / (2 * qz))
    assert rel_l2(out, ana) < 1e-6


def test_fresnel_linear_phase_shifts_beam():
    N, dx, lam, nr, dz = 128, 0.25e-6, 0.92e-6, 1.35, 20e-6
    x = (np.arange(N) - N / 2) * dx
    u = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (3e-6) ** 2).astype(np.complex128)
    f0 = 4 / (N * dx)  # exactly on the FFT grid
    tilt = np.exp(2j * np.pi * f0 * x)[:, None] * np.ones((1, N))
    phi = np.stack((tilt, np.ones((N, N))), -1)
    out = O.fresnel(u, lam, np.full(2, dx), [0, dz], phi[..., :1], nr)
    shift = lam / nr * dz * f0 / dx
    centroid = (np.abs(out) ** 2 * np.arange(N)[:, None]).sum() / (np.abs(out) ** 2).sum()
    assert abs(centroid - (N / 2 + shift)) < 1e-6


def test_fresnel_brute_force_dft_matrix():
    rng = np.random.default_rng(7)
    N, dx, lam, nr = 8, 0.5e-6, 0.92e-6, 1.35
    u = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    phi = np.exp(1j * rng.uniform(-2, 2, (N, N, 2)))
    z = [0, 1.5e-6, 4e-6]
    out, saved = O.fresnel(u, lam, np.full(3, dx), z, phi, nr, return_all=True)
    F = np.exp(-2j * np.pi * np.outer(np.arange(N), np.arange(N)) / N)
    m = np.arange(N) - N / 2
    f = m / (N * dx)
    v = u * phi[..., 0]
    for step, dzs in enumerate(np.diff(z)):
        Qc = np.exp(-1j * np.pi * lam / nr * dzs * (f[:, None] ** 2 + f[None, :] ** 2))
        # centred DFT: shift index n -> n - N/2 in space and frequency
        S = np.roll(np.eye(N), N // 2, axis=0)
        Fc = S @ F @ S
        v = np.linalg.inv(Fc) @ (Qc * (Fc @ v @ Fc.T)) @ np.linalg.inv(Fc.T)
        if step + 1 < 2:
            v = v * phi[..., step + 1]
    np.testing.assert_allclose(out, v, atol=1e-12)


def test_shift_free_form_equals_shifted_form():
    rng = np.random.default_rng(2)
    for N in (16, 14, 36):
        u = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        m = np.arange(N) - N / 2
        Qc = np.exp(-0.3j * (m[:, None] ** 2 + m[None, :] ** 2))
        a = np.fft.ifftshift(np.fft.ifft2(np.fft.ifftshift(Qc * np.fft.fftshift(np.fft.fft2(
            np.fft.fftshift(u))))))
        b = np.fft.ifft2(np.fft.ifftshift(Qc) * np.fft.fft2(u))
        np.testing.assert_allclose(a, b, atol=1e-12)


# ----------------------------------------------------------------- v4
def test_v4_matches_reference_and_reproduces_nodes(golden):
    p, v, q = golden["v4_p"], golden["v4_v"], golden["v4_q"]
    out = O.v4_interpolate(p, v, q, chunk=7)
    np.testing.assert_allclose(out, golden["v4_out"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(out[:5], v[:5], rtol=1e-9, atol=1e-9)
    single = O.v4_interpolate([[3, 4]], [[2, 7]], np.zeros((5, 6, 2)))
    np.testing.assert_array_equal(single, np.broadcast_to([2, 7], (5, 6, 2)))
    assert O.v4_interpolate(p, v, np.empty((0, 2))).shape == (0, 3)
    for bad in (0, -1):
        with pytest.raises(ValueError):
            O.v4_interpolate(p, v, q, chunk=bad)


# ----------------------------------------------------------------- pupil
def test_pupil_matches_reference(golden):
    kw = dict(volume_um=(32, 32, 20), depth_um=10, psf_um=(4, 4, 4), tail_um=5, vres=1)
    zc = _cfg(**kw, zernike_weights=tuple(golden["pupil_scene_zernike_w"]))
    assert rel_l2(O.pupil(zc), golden["pupil_scene_zernike"]) < 1e-5
    plain = _cfg(**kw)
    cp = golden["pupil_custom_phase"]
    assert rel_l2(O.pupil(plain, custom_phase=cp, phase_units="waves"),
                  golden["pupil_scene_custom"]) < 1e-5
    assert rel_l2(O.pupil(plain, custom_phase=cp, phase_units="radians",
                          phase_aperture="excitation"), golden["pupil_scene_custom_excit"]) < 1e-5
    a = O.pupil(plain, custom_phase=cp, phase_units="waves")
    b = O.pupil(plain, custom_phase=cp * 2 * np.pi, phase_units="radians")
    np.testing.assert_allclose(a, b, atol=3e-7)


def test_zernike_orthonormal_on_unit_disk():
    n = 400
    x = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y = np.meshgrid(x, x, indexing="ij")
    inside = X ** 2 + Y ** 2 < 1
    area = (2 / n) ** 2
    Z = [O.zernike(j, X, Y)[inside] for j in range(1, 12)]
    G = np.array([[np.sum(a * b) * area / np.pi for b in Z] for a in Z])
    np.testing.assert_allclose(G, np.eye(len(Z)), atol=2e-2)
    assert [(O.zidx(j, 1), O.zidx(j, 2)) for j in range(1, 8)] == [
        (0, 0), (1, 1), (1, 1), (2, 0), (2, 2), (2, 2), (3, 1)]


# ----------------------------------------------------------------- full propagate
def _case_cfg(kw):
    return _cfg(**{k: tuple(v) if isinstance(v, list) else v for k, v in kw.items()})


@pytest.mark.parametrize("name", ["scene2p", "scene_tf", "scene1p", "partial"])
def test_propagate_matches_reference(golden, name):
    meta = json.loads((GOLD / "propagate_cases.json").read_text())[name]
    cfg = _case_cfg(meta["config"])
    raw = golden_raw_fields(O.field_shapes(cfg), seed=meta["seed"])
    pts = O.locations(cfg.volume_um, meta["count"])
    res = O.propagate(cfg, raw, pts)
    pre = f"prop_{name}_"
    # the reference normalises its base beam with a float32 norm of thread-dependent
    # accumulation; rescale by the recorded ratio (psf ~ norm^(-2 order)).
    order = O.excitation_order(cfg)
    f = (meta["base_norm_f64"] / meta["base_norm_ref"]) ** (2 * order)
    assert abs(f - 1) < 2e-4
    for key in ("psf_3d", "psf", "psf_raw", "mask_raw", "energies"):
        assert rel_l2(res[key] * f, golden[pre + key]) < 2e-6, key
    for t in ("psfT", "psfB"):
        for key in ("convmask", "weight", "mask_raw"):
            assert rel_l2(res[t][key] * f, golden[f"{pre}{t}_{key}"]) < 2e-6, (t, key)
        assert rel_l2(res[t]["psfZ"], golden[f"{pre}{t}_psfZ"]) < 2e-6


def test_propagate_demo_custom_phase_matches_reference(golden):
    cfg = _cfg(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)
    raw = [np.zeros(s, np.float32) for s in O.field_shapes(cfg)]
    res = O.propagate(cfg, raw, O.locations(cfg.volume_um, 3), uin=golden["demo_uin"])
    nref, n64 = golden["demo_base_norms"]
    f = (n64 / nref) ** 4
    for key in ("psf_3d", "psf", "psf_raw", "mask_raw", "energies"):
        assert rel_l2(res[key] * f, golden[f"demo_{key}"]) < 2e-6, key
    for t in ("psfT", "psfB"):
        for key in ("convmask", "weight", "mask_raw"):
            assert rel_l2(res[t][key] * f, golden[f"demo_{t}_{key}"]) < 2e-6, (t, key)
        assert rel_l2(res[t]["psfZ"], golden[f"demo_{t}_psfZ"]) < 2e-6
    # energy contract: psf_raw.sum() equals the mean of the v4 mask at one centre site
    one = O.propagate(cfg, raw, O.locations(cfg.volume_um, 1), uin=golden["demo_uin"])
    assert abs(one["psf_raw"].sum() - one["mask_mean"]) < 1e-5 * abs(one["mask_mean"])
    np.testing.assert_allclose(one["mask"], 1)


def test_fp64_oracle_close_to_fp32_oracle(golden):
    meta = json.loads((GOLD / "propagate_cases.json").read_text())["scene2p"]
    cfg = _case_cfg(meta["config"])
    raw = golden_raw_fields(O.field_shapes(cfg), seed=meta["seed"])
    pts = O.locations(cfg.volume_um, 2)
    a = O.propagate(cfg, raw, pts, dtype=np.complex64)
    b = O.propagate(cfg, [r.astype(np.float64) for r in raw], pts, dtype=np.complex128)
    # The float32 path carries the base beam's chirp (|phase| ~ 5e4 rad at the pupil)
    # at float32 resolution; that alone costs ~1e-4 in psf relative L2 against FP64.
    assert rel_l2(a["psf_3d"], b["psf_3d"]) < 3e-4
    assert rel_l2(a["energies"], b["energies"]) < 3e-5


def test_zstack_equals_direct_convolution_with_matlab_crop():
    rng = np.random.default_rng(12)
    v = rng.random((7, 6, 5))
    for kshape in ((3, 3, 3), (4, 2, 5), (1, 4, 2)):
        p = rng.random(kshape)
        m, c = rng.random((7, 6)), rng.random((7, 6))
        got = O.zstack(v, p, m, c)
        src = v * (m * c)[..., None]
        want = np.zeros_like(v)
        cx, cy, cz = (k // 2 for k in kshape)
        for x, y, z in np.ndindex(*v.shape):
            for a, b, d in np.ndindex(*kshape):
                qx, qy, qz = x + cx - a, y + cy - b, z + cz - d
                if 0 <= qx < 7 and 0 <= qy < 6 and 0 <= qz < 5:
                    want[x, y, z] += src[qx, qy, qz] * p[a, b, d]
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
