"""GPU input generation (spectral Gaussian-process fields, raw A/B/C fields with vessel
occupancy, collection mask) against the FP64 CPU restatement, on the same normal draws."""

import numpy as np
import pytest

from oracle import OracleConfig
from oracle import inputs_np as I
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SCENE = dict(volume_um=(8.0, 8.0, 6.0), depth_um=5.0, tail_um=2.0, psf_um=(4.0, 4.0, 4.0))
DEMO = dict(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)
WIDE = dict(volume_um=(256.0, 256.0, 50.0), depth_um=100.0, vres=1.0, tail_um=20.0)


def _configs(kw):
    from maomi_amd.optics import OpticalConfig
    return OpticalConfig(**kw), OracleConfig(**kw)


def _draws(seed, shapes, device):
    """The reference's draw order: after manual_seed, per field the real then the imaginary
    normal draw (optics.py:146-162 with _optical_math.py:54-56)."""
    import torch
    devs = [0] if device.startswith("cuda") else []
    with torch.random.fork_rng(devices=devs):
        torch.manual_seed(seed)
        return [(torch.randn(s, device=device, dtype=torch.float32).cpu().numpy(),
                 torch.randn(s, device=device, dtype=torch.float32).cpu().numpy())
                for s in shapes]


def _vessels(cfg, seed, density=0.03):
    """Random vessel occupancy over the padded domain, deep enough for the A/B/C extents."""
    shape = np.rint(np.array(cfg.parameters()[0].vasc_sz) * cfg.vres).astype(int)
    _, za, zb, pc, nc = I._shape_parts(OracleConfig(**{k: getattr(cfg, k) for k in
                                                       ("volume_um", "depth_um", "vres",
                                                        "tail_um", "psf_um", "prop_um")}))
    shape[2] = max(shape[2], za + zb + nc)
    return np.random.default_rng(seed).random(shape) < density


@pytest.mark.parametrize("dims,K,p,mu", [((7, 5, 3), 1, 1, 0.0), ((16, 12, 4), 2, 5, 0.0),
                                         ((33, 30, 5), 4, 5, 0.25), ((64, 48, 9), 3, 1, 0.0)])
def test_grf_from_noise_matches_oracle(gpu_lib, dims, K, p, mu):
    from maomi_amd.inputs import grf_from_noise
    rng = np.random.default_rng(sum(dims) + K)
    re = rng.standard_normal(dims).astype(np.float32)
    im = rng.standard_normal(dims).astype(np.float32)
    ls = (rng.random((K, 3)) * 3 + 0.3).astype(np.float32)
    w = (rng.random(K) + 0.1).astype(np.float32)
    print(f"[grf] dims={dims} K={K} p={p} mu={mu}", flush=True)
    got = grf_from_noise(re, im, ls, p, mu=mu, l_weights=w).cpu().numpy()
    want = I.masked_gp(re, im, ls.astype(np.float64), p, mu=mu, l_weights=w.astype(np.float64))
    assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_masked_3dgp_draws_in_reference_order(gpu_lib, device):
    import torch
    from maomi_amd._optical_math import masked_3DGP_test
    dims, ls, w = (20, 18, 6), torch.tensor([[1.0, 2.0, 0.5], [3.0, 3.0, 1.5]]), [0.7, 0.2]
    torch.manual_seed(99)
    got = masked_3DGP_test(dims, ls.to(device), 5, l_weights=torch.tensor(w, device=device),
                           device=device)
    assert got.device.type == torch.device(device).type and got.dtype == torch.float32
    (re, im), = _draws(99, [dims], device)
    want = I.masked_gp(re, im, ls.numpy().astype(np.float64), 5, l_weights=np.array(w))
    assert rel_l2(got.cpu().numpy(), want) < 1e-5


@pytest.mark.parametrize("kw", [SCENE, DEMO, WIDE], ids=["scene", "demo", "wide"])
@pytest.mark.parametrize("with_vessels", [False, True])
def test_prepare_fields_matches_oracle(gpu_lib, kw, with_vessels):
    from maomi_amd.optics import prepare_fields
    cfg, ocfg = _configs(kw)
    shapes = I.field_shapes(ocfg)
    print(f"[prepare_fields] {kw} shapes={shapes} vessels={with_vessels}", flush=True)
    vessels = _vessels(cfg, 11) if with_vessels else None
    got = prepare_fields(cfg, vessels)
    want = I.prepare_fields(ocfg, _draws(cfg.seed, shapes, "cpu"), vessels=vessels)
    for g, w, s in zip(got, want, shapes):
        assert tuple(g.shape) == tuple(s) and g.device.type == "cpu"
        assert rel_l2(g.numpy(), w) < 1e-5


def test_prepare_fields_occupancy_only_and_lazy_vessels(gpu_lib):
    # the reference's own check (tests/test_workflow.py:198-213): vessels read by slabs only
    from maomi_amd.optics import collection_mask, prepare_fields
    cfg, ocfg = _configs(SCENE)
    shape = np.rint(np.array(cfg.parameters()[0].vasc_sz) * cfg.vres).astype(int)
    shape[2] = max(shape[2], 12)
    data = np.random.default_rng(11).random(shape) < .03

    class Bounded:
        shape = data.shape

        def __array__(self, *args, **kwargs):
            raise AssertionError("Vessels must be read as slabs")

        def __getitem__(self, k):
            return data[k]

    a = prepare_fields(cfg, data, scatter=False)
    b = prepare_fields(cfg, Bounded(), scatter=False)
    want = I.prepare_fields(ocfg, None, vessels=data, scatter=False)
    for x, y, w in zip(a, b, want):
        np.testing.assert_array_equal(x.numpy(), y.numpy())
        np.testing.assert_allclose(x.numpy(), w, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(collection_mask(data, cfg), collection_mask(Bounded(), cfg))


@pytest.mark.parametrize("kw,grid", [(SCENE, None), (DEMO, None), (WIDE, None),
                                     (dict(volume_um=(9.0, 7.0, 6.0), depth_um=9.0, tail_um=2.0,
                                           psf_um=(4, 4, 4)), (21, 19, 12))],
                         ids=["scene", "demo", "wide", "odd"])
def test_collection_mask_matches_oracle(gpu_lib, kw, grid):
    from maomi_amd.optics import collection_mask
    cfg, ocfg = _configs(kw)
    rng = np.random.default_rng(5)
    v = rng.random(grid) < 0.1 if grid else _vessels(cfg, 5, 0.05)
    print(f"[collection_mask] {kw} vessels={v.shape}", flush=True)
    got = collection_mask(v, cfg)
    want = I.collection_mask(v, ocfg)
    assert got.dtype == np.float32 and got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)
    assert got.min() < 1.0
