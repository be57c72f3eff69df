"""Write golden vectors for the optical-stage oracle from the reference itself.

Run once in a container that holds the reference sources at /root/reference:

    python tests/golden/make_golden.py

It imports the reference's CPU modules (maomi.optics, maomi.numerics,
maomi._propagation, maomi.fixed's pure-numpy helpers) and stores their outputs
on small inputs as .npz fixtures next to this script.  The fixtures are data:
inputs and expected outputs.  Nothing here is used at test time except the
.npz files; the reference never travels with the repository.

Inputs that are random are produced by seeded NumPy generators recorded in the
fixture (``golden_raw_fields``) so the tests regenerate them exactly.
"""

import json
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(os.environ.get("MAOMI_REFERENCE", "/root/reference"))


def golden_raw_fields(shapes, seed, amplitude=0.6):
    """Seeded smooth random raw index fields (units: delta-n / n_diff summed per slab)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    out = []
    for s in shapes:
        a = gaussian_filter(rng.standard_normal(s), sigma=(1.5, 1.5, 0.7), mode="wrap")
        out.append((amplitude * a / a.std()).astype(np.float32))
    return out


def main():
    sys.path.insert(0, str(REF))
    import torch
    torch.set_num_threads(8)
    from maomi.optics import OpticalConfig, pupil, propagate, locations
    from maomi import numerics, _propagation, excitation
    from maomi.fixed import fixed_locations, field_shapes, assemble

    out = {}

    # ---- geometry on several configurations
    cfgs = {
        "demo": dict(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0),
        "scene": dict(volume_um=(32, 32, 20), depth_um=10, psf_um=(4, 4, 4), tail_um=5, vres=1),
        "default": dict(),
        "partial": dict(volume_um=(12.0, 12.0, 7.0), depth_um=13.0, vres=1.0, tail_um=7.0,
                        psf_um=(4.0, 4.0, 4.0)),
        "nocrop": dict(volume_um=(8.0, 8.0, 4.0), depth_um=6.0, vres=1.0, tail_um=4.0,
                       psf_um=(3.0, 3.0, 4.0), propcrop=False),
        "odd": dict(volume_um=(8.0, 8.0, 4.0), depth_um=6.0, vres=0.75, tail_um=8.0,
                    psf_um=(4.0, 4.0, 4.0), ss=2),
        "odd_ok": dict(volume_um=(8.0, 8.0, 4.0), depth_um=6.0, vres=0.75, tail_um=8.0,
                       psf_um=(2.0, 2.0, 4.0), ss=2),
        "notail": dict(volume_um=(8.0, 8.0, 4.0), depth_um=2.0, vres=1.0, tail_um=30.0,
                       psf_um=(4.0, 4.0, 4.0)),
    }
    geo = {}
    for name, kw in cfgs.items():
        cfg = OpticalConfig(**kw)
        shapes = field_shapes(cfg)
        raw = [np.zeros(s, np.float32) for s in shapes]
        try:
            g = _propagation.geometry(cfg, raw)
        except ValueError as exc:
            geo[name] = dict(config=kw, shapes=[list(map(int, s)) for s in shapes],
                             error=str(exc))
            continue
        geo[name] = dict(
            config=kw, shapes=[list(map(int, s)) for s in shapes], n=g.n, n2=g.n2,
            fl=float(g.fl), d1=float(g.d1), d2=float(g.d2), wvl=float(g.wvl), nr=float(g.nr),
            z=[list(map(float, z)) for z in g.z], cut=g.cut,
            focal_shape=list(map(int, g.focal_shape)),
            crop=[int(g.crop.start), int(g.crop.stop)],
            focal_crop=[[int(s.start), int(s.stop)] for s in g.focal_crop],
        )
    (HERE / "geometry.json").write_text(json.dumps(geo, indent=1))

    # ---- edge window, resize, translate
    for size in (8, 14, 36):
        out[f"edge_{size}"] = _propagation.edge_apodization(size, "cpu", torch.float32).numpy()
    rng = np.random.default_rng(5)
    a = rng.standard_normal((9, 7, 3)).astype(np.float32)
    out["resize_in"] = a
    for s, tag in ((2, "2"), (0.5, "half"), (1 / 3, "third")):
        out[f"resize_{tag}"] = numerics.resize_xy(torch.from_numpy(a), s).numpy()
    b = rng.standard_normal((8, 6, 2)).astype(np.float32)
    out["translate_in"] = b
    for s, tag in ((0.5, "half"), (1.0, "one"), (1.5, "onehalf")):
        out[f"translate_{tag}"] = numerics.translate_xy(torch.from_numpy(b), s).numpy()

    # ---- fresnel: constant pitch multi-screen and one scaled step
    u = (rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))).astype(np.complex64)
    phi = np.exp(1j * rng.uniform(-1, 1, (32, 32, 3))).astype(np.complex64)
    z = np.array([0, 2e-6, 4e-6, 7e-6])
    dx = np.full(4, 0.5e-6)
    o, saved = numerics.fresnel(torch.from_numpy(u), np.float32(0.92e-6), dx, z,
                                torch.from_numpy(phi), np.float32(1.35), return_all=True,
                                crop=(slice(8, 24), slice(4, 28)), save_from=1)
    out.update(fresnel_u=u, fresnel_phi=phi, fresnel_z=z, fresnel_dx=dx,
               fresnel_out=o.numpy(), fresnel_saved=saved.numpy())
    sg = _propagation.edge_apodization(32, "cpu", torch.float32)
    so = numerics.fresnel(torch.from_numpy(u), np.float32(0.92e-6), [np.float32(4.4e-5),
                          np.float32(0.25e-6)], [0, np.float32(4.48e-3)],
                          torch.stack((sg, sg), -1), np.float32(1.35))
    out["fresnel_scaled_out"] = so.numpy()

    # ---- v4
    p = rng.uniform(-20, 20, (23, 2))
    v = rng.uniform(0.2, 2, (23, 3))
    q = np.concatenate((p[:5], rng.uniform(-25, 25, (40, 2))))
    out.update(v4_p=p, v4_v=v, v4_q=q, v4_out=numerics.v4_interpolate(p, v, q, backend="cpu"))

    # ---- pupils
    scene = OpticalConfig(**cfgs["scene"], zernike_weights=tuple(
        np.r_[0, 0, 0, rng.normal(0, 0.05, 12)] / (2 * np.pi * 1.35)))
    out["pupil_scene_zernike"] = pupil(scene).numpy()
    out["pupil_scene_zernike_w"] = np.asarray(scene.zernike_weights)
    cp = rng.normal(size=(17, 19)).astype(np.float32) * 0.1
    plain = OpticalConfig(**cfgs["scene"])
    out["pupil_custom_phase"] = cp
    out["pupil_scene_custom"] = pupil(plain, custom_phase=cp, phase_units="waves").numpy()
    out["pupil_scene_custom_excit"] = pupil(plain, custom_phase=cp, phase_units="radians",
                                            phase_aperture="excitation").numpy()

    # ---- full propagate at the scene geometry with seeded scattering fields
    cases = {
        "scene2p": (dict(cfgs["scene"]), 3, 101),
        "scene_tf": (dict(cfgs["scene"], scaling="temporal-focusing"), 2, 102),
        "scene1p": (dict(cfgs["scene"], scaling="one-photon"), 2, 103),
        "partial": (dict(cfgs["partial"]), 2, 104),
    }
    def base_norms(cfg, raw, uin=None):
        # The reference normalises its base beam with a float32 torch.linalg.vector_norm
        # whose accumulation order depends on the thread count; record that value and the
        # FP64 norm of the same base field so tests can separate this scale from the rest.
        g = _propagation.geometry(cfg, raw)
        u = pupil(cfg) if uin is None else uin
        sgn = _propagation.edge_apodization(g.n, "cpu", torch.float32)
        dz0 = np.float32(g.fl - (cfg.depth_um + cfg.volume_um[2] / 2) * 1e-6)
        b = numerics.fresnel(u, g.wvl, [g.d1, g.d2], [0, dz0], torch.stack((sgn, sgn), -1), g.nr)
        n64 = float(np.sqrt(np.sum(np.abs(b.numpy().astype(np.complex128)) ** 2)))
        return float(torch.linalg.vector_norm(b)), n64

    meta = {}
    for name, (kw, count, seed) in cases.items():
        cfg = OpticalConfig(**kw)
        raw = golden_raw_fields(field_shapes(cfg), seed=seed)
        pts = locations(cfg.volume_um, count)
        res = propagate(cfg, [torch.from_numpy(r) for r in raw], pts)
        nref, n64 = base_norms(cfg, [torch.from_numpy(r) for r in raw])
        meta[name] = dict(config=kw, count=count, seed=seed, base_norm_ref=nref,
                          base_norm_f64=n64)
        for key in ("psf_3d", "psf", "psf_raw", "mask_raw", "energies"):
            out[f"prop_{name}_{key}"] = np.asarray(res[key])
        out[f"prop_{name}_mask_mean"] = np.asarray(res["mask_mean"])
        for t in ("psfT", "psfB"):
            for key in ("convmask", "weight", "psfZ", "mask_raw", "mask_mean"):
                out[f"prop_{name}_{t}_{key}"] = np.asarray(res[t][key])
    (HERE / "propagate_cases.json").write_text(json.dumps(meta, indent=1))

    # ---- demo geometry, no scattering, custom phase: the reference's own GPU parity inputs
    demo = OpticalConfig(**cfgs["demo"])
    raw = [np.zeros(s, np.float32) for s in field_shapes(demo)]
    phase = np.zeros((17, 17), np.float32)
    phase[5:11, 5:11] = 0.2
    u0 = pupil(demo, custom_phase=phase)
    pts = locations(demo.volume_um, 3)
    res = propagate(demo, [torch.from_numpy(r) for r in raw], pts, uin=u0)
    out["demo_uin"] = u0.numpy()
    out["demo_base_norms"] = np.array(base_norms(demo, [torch.from_numpy(r) for r in raw], u0))
    for key in ("psf_3d", "psf", "psf_raw", "mask_raw", "energies"):
        out[f"demo_{key}"] = np.asarray(res[key])
    for t in ("psfT", "psfB"):
        for key in ("convmask", "weight", "psfZ", "mask_raw"):
            out[f"demo_{t}_{key}"] = np.asarray(res[t][key])

    # ---- lattice helpers
    out["fixed_demo_65"] = fixed_locations(demo)
    out["fixed_scene_3"] = fixed_locations(plain, 3)
    out["axial_tf_focus"] = excitation.axial_profile(OpticalConfig(scaling="temporal-focusing"), 20)
    out["axial_tf_top"] = excitation.axial_profile(OpticalConfig(scaling="temporal-focusing"), 7,
                                                   "top")
    out["axial_tf_bottom"] = excitation.axial_profile(
        OpticalConfig(scaling="temporal-focusing"), 5, "bottom")

    np.savez_compressed(HERE / "reference_vectors.npz", **out)
    print("wrote", HERE / "reference_vectors.npz",
          os.path.getsize(HERE / "reference_vectors.npz") // 1024, "KiB")


if __name__ == "__main__":
    main()
