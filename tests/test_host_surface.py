"""Host-side call surface of the package on the CPU: geometry, lattices, tables, pupil mirror."""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from tests.conftest import rel_l2

GOLD = Path(__file__).resolve().parent / "golden"


def test_geometry_and_shapes_match_reference():
    from maomi_amd._propagation import geometry, center_slice
    from maomi_amd.fixed import field_shapes
    from maomi_amd.optics import OpticalConfig
    geo = json.loads((GOLD / "geometry.json").read_text())
    for name, g in geo.items():
        cfg = OpticalConfig(**{k: tuple(v) if isinstance(v, list) else v
                               for k, v in g["config"].items()})
        shapes = field_shapes(cfg)
        assert [list(s) for s in shapes] == g["shapes"], name
        raw = [np.zeros(s, np.float32) for s in shapes]
        if "error" in g:
            with pytest.raises(ValueError):
                geometry(cfg, raw)
            continue
        mine = geometry(cfg, raw)
        assert (mine.n, mine.n2, mine.cut) == (g["n"], g["n2"], g["cut"])
        assert [float(mine.d1), float(mine.d2), float(mine.fl)] == [g["d1"], g["d2"], g["fl"]]
        for a, b in zip(mine.z, g["z"]):
            np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError, match="Incompatible"):
        center_slice(9, 4)


def test_lattice_and_rejections(golden):
    from maomi_amd.fixed import fixed_locations
    from maomi_amd.optics import OpticalConfig, locations
    demo = OpticalConfig(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)
    np.testing.assert_array_equal(fixed_locations(demo), golden["fixed_demo_65"])
    assert fixed_locations(demo).shape == (4225, 2)
    for bad in (1, 3.5, True):
        with pytest.raises(ValueError):
            fixed_locations(demo, bad)
    with pytest.raises(ValueError):
        fixed_locations(OpticalConfig(volume_um=(4.0, 4.0, 4.0)))
    np.testing.assert_array_equal(locations((16, 16), 1), [[8, 8]])


def test_resampling_helpers_match_reference(golden):
    from maomi_amd.numerics import resize_xy, translate_xy, resample_matrix, resize_xy_region
    a = torch.from_numpy(golden["resize_in"])
    for s, tag in ((2, "2"), (0.5, "half"), (1 / 3, "third")):
        np.testing.assert_allclose(resize_xy(a, s).numpy(), golden[f"resize_{tag}"], atol=2e-6)
    full = resize_xy(a, 2)
    part = resize_xy_region(a, 2, (3, 2), (6, 5))
    np.testing.assert_allclose(part.numpy(), full[3:9, 2:7].numpy(), atol=2e-6)
    b = torch.from_numpy(golden["translate_in"])
    for s, tag in ((0.5, "half"), (1.0, "one"), (1.5, "onehalf")):
        np.testing.assert_allclose(translate_xy(b, s).numpy(), golden[f"translate_{tag}"], atol=2e-6)
    np.testing.assert_allclose(resample_matrix(20, 2), O.resample_matrix(20, 2), atol=1e-15)


def test_torch_pupil_mirror_matches_reference(golden):
    from maomi_amd.optics import OpticalConfig, pupil
    kw = dict(volume_um=(32, 32, 20), depth_um=10, psf_um=(4, 4, 4), tail_um=5, vres=1)
    zc = OpticalConfig(**kw, zernike_weights=tuple(golden["pupil_scene_zernike_w"]))
    assert rel_l2(pupil(zc).numpy(), golden["pupil_scene_zernike"]) < 1e-5
    cp = golden["pupil_custom_phase"]
    plain = OpticalConfig(**kw)
    got = pupil(plain, custom_phase=cp, phase_units="waves").numpy()
    assert rel_l2(got, golden["pupil_scene_custom"]) < 1e-5
    with pytest.raises(ValueError):
        pupil(plain, custom_phase=cp, phase_units="degrees")


def test_excitation_and_config_errors():
    from maomi_amd.excitation import axial_profile, excitation_order, temporal_profile
    from maomi_amd.optics import OpticalConfig
    assert excitation_order(OpticalConfig(scaling="three-photon")) == 3
    with pytest.raises(ValueError, match="Unknown excitation"):
        excitation_order(OpticalConfig(scaling="x"))
    with pytest.raises(ValueError, match="temporal"):
        OpticalConfig(temporal_length_um=0).parameters()
    with pytest.raises(ValueError):
        temporal_profile(0, 10, 1)
    tf = OpticalConfig(scaling="temporal-focusing")
    np.testing.assert_allclose(axial_profile(tf, 9, "top"), O.axial_profile(tf, 9, "top"))


def test_v4_cpu_backend_and_argument_errors(golden):
    from maomi_amd.numerics import v4_interpolate
    p, v, q = golden["v4_p"], golden["v4_v"], golden["v4_q"]
    np.testing.assert_allclose(v4_interpolate(p, v, q, chunk=5, backend="cpu"), golden["v4_out"],
                               rtol=1e-9, atol=1e-10)
    for kw in ({"backend": "x"}, {"chunk": 0}):
        with pytest.raises(ValueError):
            v4_interpolate(p, v, q, **kw)
    out = v4_interpolate([[3, 4]], [[2, 7]], np.zeros((5, 6, 2)))
    np.testing.assert_array_equal(out, np.broadcast_to([2, 7], (5, 6, 2)))


def test_plain_workload_inputs_are_seeded():
    from maomi_amd.workloads import CONFIG1, CONFIG2, index_volume, site_offsets, zernike_weights
    a, b = index_volume(CONFIG1), index_volume(CONFIG1)
    np.testing.assert_array_equal(a, b)
    assert a.shape == (256, 256, 16) and abs(a.std() - 0.02) < 1e-6
    assert site_offsets(CONFIG2).shape == (4225, 2) and site_offsets(CONFIG2).max() == 512
    w = zernike_weights(CONFIG2)
    assert len(w) == 15 and w[:3] == (0.0, 0.0, 0.0) and any(w[3:])
    assert CONFIG2.bytes_per_site() == (20 * 60 + 12) * 256 * 256
