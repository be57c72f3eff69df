"""GPU arithmetic of the 2-D PSF contracts and the optical stage as a 2-D provider.

resample_psf and pupil_to_intensity (FP64 on the GPU) against the oracle's restatement, which
calls the reference's own library routines (scipy map_coordinates, numpy FFT): bar 1e-13
relative (FP64, a different FFT and summation order).  The stage provider against the oracle's
propagate of the same sites: the reference's GPU-vs-CPU bar, 2e-5.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from oracle import psf_np as OP
from tests.conftest import rel_l2
from tests.golden.make_golden import golden_raw_fields

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("src_shape,src_pitch,pitch,shape", [
    ((16, 16), (0.5, 0.5), (0.5, 0.5), (16, 16)),
    ((16, 15), (0.5, 0.4), (0.4, 0.4), (16, 15)),   # rounded edge coordinates fall outside
    ((21, 17), (0.25, 0.3), (0.35, 0.2), (30, 9)),
    ((9, 9), (0.5, 0.5), (0.125, 0.125), (40, 40)),
    ((33, 33), (0.1, 0.1), (0.9, 0.7), (5, 6)),
])
def test_resample_psf_matches_oracle(gpu_lib, src_shape, src_pitch, pitch, shape):
    from maomi_amd.psf import SampledPSF, resample_psf
    rng = np.random.default_rng(sum(src_shape) + sum(shape))
    k = SampledPSF.normalized(rng.random(src_shape) + 0.01, src_pitch, 920.0, {"id": 1})
    print(f"[resample] {src_shape} @ {src_pitch} -> {shape} @ {pitch}", flush=True)
    got = resample_psf(k, pitch, shape)
    want, mass = OP.resample_linear(k.intensity, pitch, src_pitch, shape)
    assert got.intensity.shape == shape and got.pixel_size_um == tuple(pitch)
    np.testing.assert_allclose(got.intensity, want, rtol=1e-13, atol=1e-16)
    np.testing.assert_array_equal(got.intensity == 0, want == 0)  # identical support
    assert abs(got.provenance["pre_normalization_mass_estimate"] - mass) <= 1e-13 * mass
    assert got.provenance["source"] == {"id": 1}
    assert got.provenance["source_shape"] == list(src_shape)


@pytest.mark.parametrize("pshape,n,blocked,amp", [((8, 8), 32, False, False),
                                                  ((7, 9), 33, True, False),
                                                  ((31, 30), 64, True, True),
                                                  ((64, 64), 256, False, True),
                                                  ((5, 5), 5, False, False)])
def test_pupil_to_intensity_matches_oracle(gpu_lib, pshape, n, blocked, amp):
    from maomi_amd.psf import pupil_to_intensity
    rng = np.random.default_rng(n)
    opd = rng.standard_normal(pshape) * 0.3
    if blocked:
        opd[rng.random(pshape) < 0.2] = np.nan
    a = rng.random(pshape) * (rng.random(pshape) > 0.1) if amp else None
    print(f"[pupil] {pshape} in {n}", flush=True)
    got = pupil_to_intensity(opd, n, a)
    want = OP.pupil_to_intensity(opd, n, a)
    assert got.shape == (n, n)
    assert abs(got.sum() - 1) < 1e-12
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-13 * want.max())


def test_optical_stage_provider_samples_a_psf_field(gpu_lib):
    """PSFField.from_provider over a 3 x 3 grid with the stage provider: one persistent
    launch for the grid; each kernel is the oracle's focal PSF of that site reduced over z."""
    from maomi_amd.optics import OpticalConfig
    from maomi_amd.providers import OpticalStagePSF
    from maomi_amd.psf import PSFField
    meta = json.loads((GOLD / "propagate_cases.json").read_text())["scene2p"]
    kw = {k: tuple(v) if isinstance(v, list) else v for k, v in meta["config"].items()}
    cfg, ocfg = OpticalConfig(**kw), O.OracleConfig(**kw)
    raw = golden_raw_fields(O.field_shapes(ocfg), seed=meta["seed"])
    xs, ys = np.array([0.0, 16.0, 32.0]), np.array([0.0, 8.0, 24.0])
    print("[psf provider] scene2p, 3 x 3 grid", flush=True)
    for projection in ("sum", "focus"):
        prov = OpticalStagePSF(cfg, raw, projection=projection)
        f = PSFField.from_provider(prov, xs, ys)
        assert prov.launches == 1
        assert f.pixel_size_um == (1.0, 1.0) and f.wavelength_nm == 920.0
        pts = np.stack(np.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
        ref = np.abs(O.propagate(ocfg, raw, pts)["psf_3d"].astype(np.float64))
        ref = ref.sum(-1) if projection == "sum" else ref[..., ref.shape[-1] // 2]
        ref /= ref.sum((1, 2), keepdims=True)
        assert rel_l2(f.intensity.reshape(ref.shape), ref) < 2e-5
        # a node outside the prefetched grid is evaluated on its own launch
        one = prov(4.0, 4.0)
        assert prov.launches == 2 and one.intensity.shape == f.intensity.shape[2:]
