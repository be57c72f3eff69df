"""GPU local-PSF FFT plan against the FP64 CPU restatement (which is pinned to a direct sum)."""

import itertools

import numpy as np
import pytest

from oracle import local_fft_np as L
from tests.conftest import rel_l2
from tests.test_oracle_local_fft import case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("plane,mode", list(itertools.product((False, True),
                                                              ("detection", "excitation", "native"))))
def test_plan_matches_oracle(gpu_lib, plane, mode):
    from maomi_amd._local_fft import LocalFFTPlan
    det, corr = mode == "detection", mode == "excitation"
    for seed, shape, nodes, kshape in ((7, (13, 11, 4), (3, 3), (5, 3, 3)),
                                       (8, (40, 36, 6), (4, 3), (7, 9, 5)),
                                       (9, (10, 9, 3), (2, 3), (4, 2, 2))):
        kernels, w, src = case(seed, shape=shape, nodes=nodes, kshape=kshape, plane=plane)
        print(f"[local fft] {mode} plane={plane} shape={shape} kernel={kernels.shape}", flush=True)
        plan = LocalFFTPlan(shape, kernels, w, plane=plane, detection=det, correlate=corr,
                            device="cuda", batch_size=3)
        got = plan(src).cpu().numpy()
        want = L.local_fft(shape, kernels, w, src, plane=plane, detection=det, correlate=corr)
        assert got.shape == want.shape
        assert rel_l2(got, want) < 1e-12


def test_cache_statistics_and_clear(gpu_lib):
    from maomi_amd._local_fft import LocalFFTPlan
    kernels, w, src = case(11, shape=(24, 20, 4), nodes=(3, 3), kshape=(5, 5, 3))
    plan = LocalFFTPlan(src.shape, kernels, w, plane=False, detection=False, correlate=True,
                        device="cuda", batch_size=4)
    a = plan(src)
    builds = plan.stats["kernel_fft_builds"]
    assert builds > 0 and plan.stats["cache_hits"] == 0 and plan.cache_bytes > 0
    b = plan(src)
    assert plan.stats["kernel_fft_builds"] == builds and plan.stats["cache_hits"] == builds
    assert float((a - b).abs().max()) < 1e-12 * float(a.abs().max())
    plan.clear()
    assert plan.cache_bytes == 0 and not plan.cache
    tiny = LocalFFTPlan(src.shape, kernels, w, plane=False, detection=False, correlate=True,
                        device="cuda", batch_size=4, cache_mb=0)
    tiny(src)
    assert tiny.cache_bytes == 0 and tiny.stats["kernel_fft_builds"] == builds


def test_rejections(gpu_lib):
    from maomi_amd._local_fft import LocalFFTPlan
    kernels, w, src = case(12)
    with pytest.raises(ValueError):
        LocalFFTPlan(src.shape, kernels, w, plane=False, detection=False, correlate=True,
                     device="cpu")
    with pytest.raises(ValueError):
        LocalFFTPlan(src.shape, kernels, w, plane=True, detection=False, correlate=True,
                     device="cuda")  # kernel depth 3 != source depth 4
    plan = LocalFFTPlan(src.shape, kernels, w, plane=False, detection=False, correlate=True,
                        device="cuda")
    with pytest.raises(ValueError):
        plan(src[:, :, :2])
