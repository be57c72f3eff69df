"""GPU parity of the alternative input pupils (maomi_pupil_field) against the oracle
(oracle/pupils_np.py, pinned in tests/test_oracle_pupils.py).

Both sides use the same float32 scalars and the same float32 operation order; they differ only
in the libraries' exp, sin and cos (within 2 ulp each), so the bar is 1e-6 of the field's peak
(float32 resolution at the peak is 6e-8).  Ring supports are pure comparisons of equal float32
values and must match exactly."""

import numpy as np
import pytest

import oracle as O
from oracle import pupils_np as PO

pytestmark = pytest.mark.gpu
DEMO = dict(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)

CASES = [
    ("pupil", dict()),
    ("pupil", dict(kind="cylindrical", cylscale=(1.6, 0.55))),
    *[("pupil", dict(kind="bessel", profile=p, length_um=40.0))
      for p in ("gaussian", "uniform", "posax", "negax")],
    ("pair", dict()),
    ("pair", dict(kind="bessel", profile="posax", theta_deg=-30.0, separation_um=4.0)),
    ("pair", dict(kind="bessel", profile="uniform", clip_to_objective=False, beam_na=0.5,
                  theta_deg=60.0, length_um=25.0)),
    ("pair", dict(kind="bessel", profile="negax", beam_na=0.3, separation_um=9.0)),
]


def _check(got, ref, exact_support):
    got = got.cpu().numpy()
    assert got.shape == ref.shape and got.dtype == np.complex64
    peak = float(np.abs(ref).max())
    err = float(np.abs(got.astype(np.complex128) - ref).max())
    assert err <= 1e-6 * peak, (err, peak)
    if exact_support:
        np.testing.assert_array_equal(got != 0, ref != 0)


@pytest.mark.parametrize("geometry", ["default", "demo"])
@pytest.mark.parametrize("which,kw", CASES)
def test_pupil_fields_match_oracle(gpu_lib, geometry, which, kw):
    from maomi_amd import pupils as P
    from maomi_amd.optics import OpticalConfig
    ckw = {} if geometry == "default" else DEMO
    cfg, ocfg = OpticalConfig(**ckw), O.OracleConfig(**ckw)
    print(f"pupil field {which} {kw} on the {geometry} grid", flush=True)
    exact = kw.get("profile") == "uniform" or kw.get("kind") == "cylindrical"
    if which == "pupil":
        _check(P.make_pupil(cfg, **kw), PO.make_pupil(ocfg, **kw), exact)
    else:
        got, ref = P.make_vtwins_pupils(cfg, **kw), PO.make_vtwins_pupils(ocfg, **kw)
        assert set(got) == {"left", "right"}
        for label in ("left", "right"):
            _check(got[label], ref[label], exact)


def test_pair_drives_run_fixed_channels(gpu_lib):
    """A vTwINS pair as run_fixed's input_pupil mapping: one native run per beam
    (fixed.py:239-287), each the same as a run with that beam alone."""
    from maomi_amd import pupils as P
    from maomi_amd.fixed import run_fixed
    from maomi_amd.optics import OpticalConfig, prepare_fields
    cfg = OpticalConfig(**DEMO)
    raw = prepare_fields(cfg, None)
    pair = P.make_vtwins_pupils(cfg, theta_deg=16.0, separation_um=4.0)
    print("run_fixed with a vTwINS pair, 3 x 3 sites", flush=True)
    res = run_fixed(cfg, None, raw_fields=raw, input_pupil=pair, count_per_axis=3,
                    render_image=False)
    assert set(res["channels"]) == {"left", "right"}
    left = run_fixed(cfg, None, raw_fields=raw, input_pupil=pair["left"], count_per_axis=3,
                     render_image=False)
    np.testing.assert_array_equal(res["channels"]["left"]["psf"], left["psf"])
    a, b = res["channels"]["left"]["psf"], res["channels"]["right"]["psf"]
    assert np.isfinite(a).all() and np.isfinite(b).all() and a.sum() > 0 and b.sum() > 0
    assert not np.array_equal(a, b)
