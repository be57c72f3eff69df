"""Device v4 (griddata 'v4') against the reference's own vectors and a direct FP64 sum.

The evaluation uses its own FP64 logarithm (atanh series, within 2 ulp), so each Green term
may differ from NumPy's in the last bits; the bar for a query is therefore relative to the
sum of the magnitudes of its terms.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_v4_interpolate_cuda_matches_reference_vectors(gpu_lib, golden):
    from maomi_amd.numerics import v4_interpolate
    p, v, q = golden["v4_p"], golden["v4_v"], golden["v4_q"]
    got = v4_interpolate(p, v, q, backend="cuda")
    np.testing.assert_allclose(got, golden["v4_out"], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("P,Q,C,rows", [(50, 700, 1, 1), (1100, 3000, 3, 3), (300, 257, 8, 2),
                                        (4225, 5000, 1, 0)])
def test_v4_apply_matches_direct_sum(gpu_lib, P, Q, C, rows):
    import torch
    from maomi_amd import _native as nat
    rng = np.random.default_rng(P + Q)
    p = rng.random((P, 2)) * 500
    q = np.concatenate([p[: min(P, Q // 4)], rng.random((Q - min(P, Q // 4), 2)) * 520 - 10])
    q[:3] = p[0] + np.array([[1e-7, 0], [0, 3e-4], [40.0, 0]])  # tiny, small and large distances
    coef = rng.standard_normal((P, C))
    d2 = ((q[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = d2 * (0.5 * np.log(d2) - 1)
    g[d2 == 0] = 0
    ref = g @ coef
    bound = np.abs(g) @ np.abs(coef)
    dev = torch.device("cuda")
    tp, tc, tq = (torch.as_tensor(a, device=dev) for a in (p, coef, q))
    out = torch.empty((Q, C), dtype=torch.float64, device=dev)
    lib = gpu_lib
    rows = rows or int(lib.maomi_v4_apply_rows(P, Q))
    part = torch.empty(rows * Q * C, dtype=torch.float64, device=dev)
    print(f"[v4 apply] P={P} Q={Q} C={C} rows={rows}", flush=True)
    nat.check(lib.maomi_v4_apply_split_f64(nat.ptr(tp), nat.ptr(tc), nat.ptr(tq), nat.ptr(out),
                                           P, Q, C, nat.ptr(part), rows, nat.stream_handle(dev)),
              "v4 apply")
    got = out.cpu().numpy()
    assert (np.abs(got - ref) <= 1e-13 * bound + 1e-300).all()
