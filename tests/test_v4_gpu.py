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


@pytest.mark.parametrize("name", ["collinear", "three", "two", "one", "lattice"])
def test_v4_interpolate_cuda_matches_cpu_backend(gpu_lib, name):
    """Device solve routes (Cholesky of the projected system, LU fallback for collinear or
    P <= 3, single point) against the NumPy backend (numerics.py:122-138)."""
    from maomi_amd.numerics import v4_interpolate
    rng = np.random.default_rng(11)
    if name == "collinear":  # rank-2 polynomial space: the LU fallback
        p = np.stack([np.linspace(0, 10, 40), np.linspace(0, 5, 40)], 1)
    elif name == "lattice":
        t = np.arange(9) * 4.0
        xx, yy = np.meshgrid(t, t, indexing="ij")
        p = np.stack([xx.ravel(), yy.ravel()], 1)
    else:
        p = rng.random(({"three": 3, "two": 2, "one": 1}[name], 2)) * 10
    v = rng.standard_normal((len(p), 2))
    q = rng.random((17, 23, 2)) * 12 - 1
    got = v4_interpolate(p, v, q, backend="cuda")
    ref = v4_interpolate(p, v, q, backend="cpu")
    assert got.shape == ref.shape == (17, 23, 2)
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())


def test_cholesky_of_empty_matrix(gpu_lib):
    import torch
    from maomi_amd.numerics import cholesky_f64_
    assert cholesky_f64_(torch.empty((0, 0), dtype=torch.float64, device="cuda")) == 0


@pytest.mark.parametrize("case", ["config2_lattice", "shuffled_uneven_grid", "narrow_rows"])
def test_v4_apply_on_a_grid_matches_the_per_pair_kernel_and_direct_sum(gpu_lib, case):
    """maomi_v4_apply_grid_f64 (G from a table of axis distances) against the per-pair kernel
    on the same inputs and a direct FP64 sum; the sums run in another point order, so the bar
    is relative to the sum of the magnitudes of the terms."""
    import torch
    from maomi_amd.numerics import point_grid, v4_apply, v4_apply_grid
    rng = np.random.default_rng(len(case))
    if case == "config2_lattice":  # the 65 x 65 sites and the 512^2 pixel centres
        from maomi_amd.workloads import CONFIG2, site_offsets
        p = site_offsets(CONFIG2).astype(np.float64)
        qx = qy = np.arange(512) + 0.5
        C = 1
    elif case == "shuffled_uneven_grid":  # uneven spacings, points in random order, C = 3
        xs = np.sort(rng.choice(np.arange(0.0, 300.0, 0.5), 23, replace=False))
        ys = np.sort(rng.choice(np.arange(0.0, 300.0, 0.5), 17, replace=False))
        p = np.stack(np.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)[rng.permutation(23 * 17)]
        qx, qy = np.arange(40) * 7.5 + 0.25, np.arange(33) * 9.0 + 0.5
        C = 3
    else:  # one point column per block row, a query grid narrower than a block
        xs, ys = np.arange(3) * 10.0, np.arange(90) * 2.0
        p = np.stack(np.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
        qx, qy = np.arange(5) + 0.5, np.arange(7) * 3.0
        C = 2
    grid = point_grid(p)
    assert grid is not None
    coef = rng.standard_normal((len(p), C))
    dev = torch.device("cuda")
    tc = torch.as_tensor(coef, device=dev)
    q = np.stack(np.meshgrid(qx, qy, indexing="ij"), -1).reshape(-1, 2)
    out = torch.empty((len(q), C), dtype=torch.float64, device=dev)
    print(f"[v4 grid] {case} P={len(p)} Q={len(q)} C={C}", flush=True)
    assert v4_apply_grid(grid, tc, qx, qy, out, dev) is out
    ref_dev = torch.empty_like(out)
    v4_apply(torch.as_tensor(p, device=dev), tc, torch.as_tensor(q, device=dev), ref_dev, dev)
    got, per_pair = out.cpu().numpy(), ref_dev.cpu().numpy()
    if len(q) * len(p) <= 2e7:
        d2 = ((q[:, None, :] - p[None, :, :]) ** 2).sum(-1)
        with np.errstate(divide="ignore", invalid="ignore"):
            g = d2 * (0.5 * np.log(d2) - 1)
        g[d2 == 0] = 0
        bound = np.abs(g) @ np.abs(coef)
        assert (np.abs(got - g @ coef) <= 1e-13 * bound + 1e-300).all()
    else:
        bound = np.abs(per_pair).max() * len(p)
    assert (np.abs(got - per_pair) <= 1e-13 * bound + 1e-300).all()
