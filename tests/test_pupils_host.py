"""Host-side behaviour of maomi_amd.pupils that needs no GPU: the reference's argument checks
(maomi/pupils.py:7-122), the float32 grid against the oracle's, and the refusal of a CPU
device (the fields are built by maomi_pupil_field only)."""

import numpy as np
import pytest

import oracle as O
from oracle import pupils_np as PO
from maomi_amd import pupils as P
from maomi_amd.optics import OpticalConfig

DEMO = dict(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)


@pytest.mark.parametrize("kw", [{}, DEMO])
def test_grid_matches_oracle(kw):
    g, og = P._grid(OpticalConfig(**kw)), PO.pupil_grid(O.OracleConfig(**kw))
    np.testing.assert_array_equal(g.xs, og.xs)
    np.testing.assert_array_equal(g.ys, og.ys)
    for name in ("rad", "objrad", "k", "fl", "nr"):
        assert getattr(g, name) == getattr(og, name) and getattr(g, name).dtype == np.float32


def test_rejections_before_any_gpu_work():
    cfg = OpticalConfig()
    with pytest.raises(ValueError, match="single-field kinds"):
        P.make_pupil(cfg, kind="airy")
    with pytest.raises(ValueError, match="Bessel length"):
        P.make_pupil(cfg, kind="bessel", length_um=np.inf)
    with pytest.raises(ValueError, match="Unknown Bessel profile"):
        P.make_pupil(cfg, kind="bessel", profile="cosine")
    with pytest.raises(ValueError, match="cylindrical scale"):
        P.make_pupil(cfg, kind="cylindrical", cylscale=(1.0, 0.0))
    for kw in [dict(beam_na=0.0), dict(beam_na=1.35), dict(theta_deg=-180.0),
               dict(length_um=0.0), dict(theta_deg=np.nan)]:
        with pytest.raises(ValueError, match="Invalid Vtwins"):
            P.make_vtwins_pupils(cfg, **kw)
    with pytest.raises(ValueError, match="gaussian or bessel"):
        P.make_vtwins_pupils(cfg, kind="cylindrical")
    with pytest.raises(ValueError, match="Unknown Bessel profile"):
        P.make_vtwins_pupils(cfg, kind="bessel", profile="flat")
    with pytest.raises(ValueError, match="cuda device"):
        P.make_pupil(cfg, device="cpu")
