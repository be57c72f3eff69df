"""PSF containers and the provider boundary (CPU): the reference's contracts
(rana/fields.py:17-110, rana/depth.py:9-69, rana/providers.py:10-40,
rana/adapters/naomi.py:38-55)."""

import json

import numpy as np
import pytest

from maomi_amd.fields import DepthPSFBank, FullFieldPSF, PSFProvider, resolve_field
from maomi_amd.providers import native_field


def field(op="excitation", seed=0, z=True):
    rng = np.random.default_rng(seed)
    k = rng.random((2, 3, 5, 5, 4) if z else (2, 3, 5, 5)) + 0.01
    return FullFieldPSF(k, [0, 6], [0, 4, 8], (1, 1, 1) if z else (1, 1), 920, op, {"s": seed})


def test_full_field_validation_messages():
    k = np.ones((2, 2, 3, 3))
    with pytest.raises(ValueError, match="complex"):
        FullFieldPSF(k + 0j, [0, 1], [0, 1], (1, 1), 920, "excitation")
    with pytest.raises(ValueError, match="EACH XY axis"):
        FullFieldPSF(k, [0], [0, 1], (1, 1), 920, "excitation")
    with pytest.raises(ValueError, match="EACH XY axis"):
        FullFieldPSF(k, [1, 0], [0, 1], (1, 1), 920, "excitation")
    with pytest.raises(ValueError, match="nonnegative finite kernels"):
        FullFieldPSF(-k, [0, 1], [0, 1], (1, 1), 920, "excitation")
    z = k.copy()
    z[0, 1] = 0
    with pytest.raises(ValueError, match="positive finite energy"):
        FullFieldPSF(z, [0, 1], [0, 1], (1, 1), 920, "excitation")
    with pytest.raises(ValueError, match="voxel pitch"):
        FullFieldPSF(k, [0, 1], [0, 1], (1, 1, 1), 920, "excitation")
    with pytest.raises(ValueError, match="wavelength_nm"):
        FullFieldPSF(k, [0, 1], [0, 1], (1, 1), -1, "excitation")
    with pytest.raises(ValueError, match="operator"):
        FullFieldPSF(k, [0, 1], [0, 1], (1, 1), 920, "scan")
    with pytest.raises(ValueError):
        FullFieldPSF(k, [0, 1], [0, 1], (1, 1), 920, "excitation", {"bad": float("nan")})
    f = FullFieldPSF(k, [0, 1], [0, 1], (1, 1), 920, "detection")
    assert f.kernels.dtype == np.float64 and not f.kernels.flags.writeable


def test_full_field_round_trip_and_schema(tmp_path):
    f = field()
    p = tmp_path / "sub" / "field.npz"
    f.save(p)
    g = FullFieldPSF.load(p)
    np.testing.assert_array_equal(g.kernels, f.kernels)
    assert (g.voxel_size_um, g.operator, g.provenance) == (f.voxel_size_um, f.operator, f.provenance)
    with np.load(p, allow_pickle=False) as d:
        assert str(d["schema"]) == "maomi.full_field_psf.v1"
        assert json.loads(str(d["provenance"])) == {"s": 0}
    np.savez(tmp_path / "bad.npz", schema="other", kernels=f.kernels)
    with pytest.raises(ValueError, match="schema"):
        FullFieldPSF.load(tmp_path / "bad.npz")


def test_from_provider_samples_without_renormalising():
    f = FullFieldPSF.from_provider(lambda x, y: np.full((3, 3), 1 + x + 10 * y), [0, 2], [0, 1, 5],
                                   voxel_size_um=(1, 1), wavelength_nm=920, operator="detection")
    assert f.kernels[1, 2, 0, 0] == 53 and f.kernels.shape == (2, 3, 3, 3)


def test_depth_bank_checks_and_round_trip(tmp_path):
    b = DepthPSFBank([0.0, 10.0], (field(seed=1), field(seed=2)))
    assert b.operator == "excitation" and b.voxel_size_um == (1.0, 1.0, 1.0)
    b.save(tmp_path / "bank.npz")
    c = DepthPSFBank.load(tmp_path / "bank.npz")
    np.testing.assert_array_equal(c.fields[1].kernels, b.fields[1].kernels)
    assert [f.provenance for f in c.fields] == [{"s": 1}, {"s": 2}]
    with pytest.raises(ValueError, match="increasing"):
        DepthPSFBank([10.0, 0.0], (field(), field()))
    with pytest.raises(ValueError, match="XYZ-kernel"):
        DepthPSFBank([0.0, 1.0], (field(z=False), field(z=False)))
    with pytest.raises(ValueError, match="share"):
        DepthPSFBank([0.0, 1.0], (field(), field(op="detection")))
    bank = DepthPSFBank.from_provider(lambda x, y, z: np.full((2, 2, 2), 1 + z), [0, 1], [0, 1],
                                      [0.0, 3.0], voxel_size_um=(1, 1, 1), wavelength_nm=920)
    assert bank.fields[1].kernels[0, 0, 0, 0, 0] == 4 and bank.fields[1].provenance["focal_depth_um"] == 3


def test_resolve_field_protocol():
    f = field()

    class Good:
        calls = 0

        def build_field(self):
            Good.calls += 1
            return f

    class Wrong:
        def build_field(self):
            return np.ones(3)

    class Failing:
        def build_field(self):
            raise RuntimeError("solver failed")

    assert resolve_field(f) is f
    assert isinstance(Good(), PSFProvider)
    assert resolve_field(Good()) is f and Good.calls == 1
    with pytest.raises(TypeError, match="build_field"):
        resolve_field(object())
    with pytest.raises(TypeError, match="must return"):
        resolve_field(Wrong())
    with pytest.raises(RuntimeError, match="solver failed"):  # no mean-PSF fallback
        resolve_field(Failing())


def test_native_field_orders_scales_and_checks_the_lattice():
    rng = np.random.default_rng(4)
    pts = np.array([[x, y] for y in (0.0, 2.0, 4.0) for x in (1.0, 3.0)])  # y-major order
    psf3 = rng.random((6, 3, 3, 2)) + 0.1
    optics = {"points_um": pts, "psf_3d": psf3, "psf": psf3.mean(0)}
    f = native_field(optics, voxel_size_um=(0.5, 0.5, 0.5), wavelength_nm=920)
    assert f.kernels.shape == (2, 3, 3, 3, 2)
    np.testing.assert_array_equal(f.x_um, [1.0, 3.0])
    target = psf3.mean(0).sum()
    np.testing.assert_allclose(f.kernels.sum(axis=(2, 3, 4)), target)
    # node (x=3, y=2) is input row 3; its shape survives the energy scaling
    k = psf3[3] / psf3[3].sum() * target
    np.testing.assert_allclose(f.kernels[1, 1], k)
    with pytest.raises(ValueError, match="rectangular"):
        native_field({"points_um": pts[:5], "psf_3d": psf3[:5], "psf": psf3[0]},
                     voxel_size_um=(1, 1, 1))
