"""maomi_cholesky_f64 against numpy's Cholesky, and the device v4 solve against LU.

The factorisation is pinned by the textbook routine (np.linalg.cholesky) on SPD matrices of
sizes around the 64-column step, inside larger row-major buffers (lda > m, offset views), and
by its pivot report on indefinite matrices (LAPACK potrf's info convention).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(m, rng):
    a = rng.standard_normal((m, m + 8))
    return a @ a.T / m + np.eye(m) * 0.1


@pytest.mark.parametrize("m", [1, 5, 63, 64, 65, 130, 300, 1000])
def test_cholesky_matches_numpy(gpu_lib, m):
    import torch
    from maomi_amd.numerics import cholesky_f64_
    rng = np.random.default_rng(m)
    A = _spd(m, rng)
    buf = rng.standard_normal((m + 5, m + 7))
    buf[3:3 + m, 2:2 + m] = A
    t = torch.as_tensor(buf, device="cuda")
    print(f"[cholesky] m={m}", flush=True)
    assert cholesky_f64_(t[3:3 + m, 2:2 + m]) == 0
    got = t.cpu().numpy()
    L = np.linalg.cholesky(A)
    low = np.tril_indices(m)
    assert np.abs(got[3:3 + m, 2:2 + m][low] - L[low]).max() <= 1e-12 * np.abs(L).max()
    up = np.triu_indices(m, 1)
    np.testing.assert_array_equal(got[3:3 + m, 2:2 + m][up], A[up])  # upper triangle untouched
    outside = np.ones(buf.shape, bool)
    outside[3:3 + m, 2:2 + m] = False
    np.testing.assert_array_equal(got[outside], buf[outside])


@pytest.mark.parametrize("m,bad", [(10, 2), (200, 70), (130, 129)])
def test_cholesky_reports_first_bad_pivot(gpu_lib, m, bad):
    import torch
    from maomi_amd.numerics import cholesky_f64_
    d = np.arange(1.0, m + 1)
    d[bad] = -1.0
    t = torch.as_tensor(np.diag(d), device="cuda")
    assert cholesky_f64_(t) == bad + 1


def _green(p):
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = d2 * (np.log(d2) * 0.5 - 1)
    g[d2 == 0] = 0
    return g


def _points(name, rng):
    if name == "lattice":
        t = np.round(np.linspace(0, 256, 33))
        xx, yy = np.meshgrid(t, t, indexing="ij")
        return np.stack([xx.ravel(), yy.ravel()], 1)
    if name == "config2":  # the 65 x 65 site lattice of the 4225-site plain workload
        from maomi_amd.workloads import CONFIG2, site_offsets
        return site_offsets(CONFIG2).astype(np.float64)
    if name == "reference_demo":  # run_fixed's 65 x 65 lattice on the demo scene, in voxels
        from maomi_amd.fixed import fixed_locations
        from maomi_amd.optics import OpticalConfig
        cfg = OpticalConfig(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)
        return fixed_locations(cfg, 65) * cfg.vres
    return rng.random((700, 2)) * 100


@pytest.mark.parametrize("name", ["lattice", "random", "config2", "reference_demo"])
def test_device_v4_coefficients_equal_lu_solution(gpu_lib, name):
    """The positive-definite route (projection + maomi_cholesky_f64 + Schur complement) against
    LAPACK's LU, at the timed workloads' full size (4225 points) too.  The factorisation's
    pivot report must be 0: the Cholesky route, not the LU fallback, produced the result."""
    import torch
    from maomi_amd.numerics import v4_coefficients, v4_factor
    rng = np.random.default_rng(3)
    p = _points(name, rng)
    G = _green(p)
    v = rng.random((len(p), 3))
    ref = np.linalg.solve(G, v)
    print(f"[v4 solve] {name}: P={len(p)}", flush=True)
    tp, tG = torch.as_tensor(p, device="cuda"), torch.as_tensor(G, device="cuda")
    f = v4_factor(tp, tG)
    assert f is not None and f.V is not None and int(f.info) == 0
    w = v4_coefficients(tp, None, torch.as_tensor(v, device="cuda"), factor=f).cpu().numpy()
    # At 4225 lattice points cond(G) is about 1.1e8, so two backward-stable solves may differ by
    # ~cond * eps in the weights.  Checked: the normwise backward error |G w - v| / (|G| |w|)
    # against 50 eps (LAPACK's LU sits near 1e-17 here), the weights against LU, and the
    # interpolated values, which are what the masks use.
    res = np.linalg.norm(G @ w - v) / np.linalg.norm(v)
    res_lu = np.linalg.norm(G @ ref - v) / np.linalg.norm(v)
    eta = np.linalg.norm(G @ w - v) / (np.linalg.norm(G) * np.linalg.norm(w))
    err = np.linalg.norm(w - ref) / np.linalg.norm(ref)
    q = rng.random((2000, 2)) * (p.max() - p.min()) + p.min()
    Gq = _green_pairs(q, p)
    ev = np.linalg.norm(Gq @ w - Gq @ ref) / np.linalg.norm(Gq @ ref)
    print(f"  residual {res:.2e} (LU {res_lu:.2e}), backward error {eta:.2e}, weights {err:.2e}, "
          f"values {ev:.2e}", flush=True)
    assert eta <= 50 * np.finfo(np.float64).eps
    assert err < (1e-9 if len(p) < 2000 else 5e-9)
    assert ev < (1e-9 if len(p) < 2000 else 3e-9)


def _green_pairs(q, p):
    d2 = ((q[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = d2 * (np.log(d2) * 0.5 - 1)
    g[d2 == 0] = 0
    return g


def _finish_reference(Gf, V, T, v):
    """w from the device's factored system by LAPACK routines on the host: potrs (scipy's
    cho_solve) for the triangular sweeps, LU for the 3 x 3 Schur complement."""
    from scipy.linalg import cho_solve
    L = np.tril(Gf[3:, 3:])
    A11, A12 = Gf[:3, :3], Gf[:3, 3:]
    y = v - V @ (T.T @ (V.T @ v))
    X = cho_solve((L, True), np.concatenate([A12.T, y[3:]], 1))
    Z, z = X[:, :3], X[:, 3:]
    a = np.linalg.solve(A11 - A12 @ Z, y[:3] - A12 @ z)
    x = np.concatenate([a, z - Z @ a], 0)
    return x - V @ (T @ (V.T @ x))


@pytest.mark.parametrize("n,C", [(4, 1), (5, 2), (67, 3), (68, 1), (131, 13), (700, 15),
                                 (4225, 3)])
def test_v4_finish_matches_lapack_on_the_same_factor(gpu_lib, n, C):
    """maomi_v4_finish_f64 (triangular sweeps in 64-row panels, pivoted 3 x 3 Schur solve,
    projections) against LAPACK on the host applied to the device's own factorisation: panel
    edges (m = n - 3 = 1, 2, 64, 65, 128), more value columns than one call takes (C = 15 runs
    as 13 + 2), and the 4225-point lattice."""
    import torch
    from maomi_amd.numerics import v4_factor, v4_finish
    rng = np.random.default_rng(n + C)
    p = _points("config2", rng) if n == 4225 else rng.random((n, 2)) * 50
    G = _green(p)
    v = rng.standard_normal((n, C))
    tG = torch.as_tensor(G, device="cuda")
    f = v4_factor(torch.as_tensor(p, device="cuda"), tG, keep_copy=False)
    assert f is not None and int(f.info) == 0
    print(f"[v4 finish] n={n} C={C}", flush=True)
    w = v4_finish(f, torch.as_tensor(v, device="cuda")).cpu().numpy()
    ref = _finish_reference(f.G.cpu().numpy(), f.V.cpu().numpy(), f.T.cpu().numpy(), v)
    err = np.linalg.norm(w - ref) / np.linalg.norm(ref)
    eta = np.linalg.norm(G @ w - v) / (np.linalg.norm(G) * np.linalg.norm(w))
    print(f"  against LAPACK on the same factor {err:.2e}, backward error {eta:.2e}", flush=True)
    assert err < 1e-9
    assert eta <= 50 * np.finfo(np.float64).eps

