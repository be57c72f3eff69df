"""The GPU optical stage as a PSF provider: build_field() against the same field assembled from
the CPU oracle's propagation of the same lattice (bar: the reference's GPU-vs-CPU 2e-5)."""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from tests.conftest import rel_l2
from tests.golden.make_golden import golden_raw_fields

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def test_build_field_matches_oracle_lattice(gpu_lib):
    from maomi_amd.fields import FullFieldPSF, resolve_field
    from maomi_amd.optics import OpticalConfig
    from maomi_amd.providers import OpticalStageProvider, native_field
    meta = json.loads((GOLD / "propagate_cases.json").read_text())["scene2p"]
    kw = {k: tuple(v) if isinstance(v, list) else v for k, v in meta["config"].items()}
    cfg, ocfg = OpticalConfig(**kw), O.OracleConfig(**kw)
    raw = golden_raw_fields(O.field_shapes(ocfg), seed=meta["seed"])
    prov = OpticalStageProvider(cfg, None, count_per_axis=3, raw_fields=raw)
    print("[provider] scene2p, 3 x 3 lattice", flush=True)
    f = resolve_field(prov)
    assert isinstance(f, FullFieldPSF) and f.kernels.shape[:2] == (3, 3)
    pts = O.fixed_locations(ocfg, 3)
    ref = O.propagate(ocfg, raw, pts)
    g = native_field({"points_um": pts, "psf_3d": ref["psf_3d"], "psf": ref["psf"]}, config=cfg)
    np.testing.assert_array_equal(f.x_um, g.x_um)
    assert f.voxel_size_um == (1 / cfg.vres,) * 3 and f.operator == "excitation"
    assert rel_l2(f.kernels, g.kernels) < 2e-5


def test_run_fixed_with_rana_attaches_native_operator(gpu_lib):
    from maomi_amd.fixed import run_fixed
    from maomi_amd.optics import OpticalConfig
    from oracle.rana_np import Operator
    meta = json.loads((GOLD / "propagate_cases.json").read_text())["scene2p"]
    kw = {k: tuple(v) if isinstance(v, list) else v for k, v in meta["config"].items()}
    cfg, ocfg = OpticalConfig(**kw), O.OracleConfig(**kw)
    raw = golden_raw_fields(O.field_shapes(ocfg), seed=meta["seed"])
    shape = tuple(np.round(np.asarray(cfg.volume_um) * cfg.vres).astype(int))
    vol = np.random.default_rng(2).random(shape).astype(np.float32)
    print(f"[provider] run_fixed rana=True, volume {shape}", flush=True)
    res = run_fixed(cfg, vol, raw_fields=raw, count_per_axis=3, rana=True)
    op = res["rana"]
    f = op.field
    assert res["image"].shape == shape and "rana" in res["timing_seconds"]
    ref = Operator([f.kernels], f.x_um, f.y_um, f.voxel_size_um, "excitation", native=True)
    a = vol.astype(float) * (res["mask"] * res["colmask"])[..., None]
    assert rel_l2(res["image"], ref.static(a)) < 1e-10
