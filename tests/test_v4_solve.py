"""The v4 solve through the positive-definite part of the Green matrix against a plain LU
solve (host logic on CPU tensors; the GPU runs the same torch operations)."""

import numpy as np
import pytest
import torch


def _green(p):
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = d2 * (np.log(d2) * 0.5 - 1)
    g[d2 == 0] = 0
    return g


def _cases():
    rng = np.random.default_rng(0)
    t = np.round(np.linspace(0, 256, 33))
    xx, yy = np.meshgrid(t, t, indexing="ij")
    yield "lattice", np.stack([xx.ravel(), yy.ravel()], 1)
    yield "random", rng.random((400, 2)) * 100
    yield "collinear", np.stack([np.linspace(0, 10, 40), np.linspace(0, 5, 40)], 1)
    yield "three", rng.random((3, 2))
    yield "two", rng.random((2, 2))


@pytest.mark.parametrize("name,p", list(_cases()), ids=[c[0] for c in _cases()])
def test_v4_coefficients_equal_lu_solution(name, p):
    from maomi_amd.numerics import v4_coefficients, v4_solve
    rng = np.random.default_rng(len(p))
    G = _green(p)
    v = rng.random((len(p), 3))
    ref = np.linalg.solve(G, v)
    w = v4_coefficients(torch.as_tensor(p), torch.as_tensor(G.copy()), torch.as_tensor(v))
    assert np.linalg.norm(w.numpy() - ref) / np.linalg.norm(ref) < 1e-9
    if name == "collinear":  # rank-2 polynomial space: the positive-definite route declines
        assert v4_solve(torch.as_tensor(p), torch.as_tensor(G.copy()), torch.as_tensor(v)) is None


def test_green_matrix_has_three_negative_eigenvalues():
    # conditional positive definiteness of order 2: only the span of 1, x, y is negative
    t = np.round(np.linspace(0, 256, 17))
    xx, yy = np.meshgrid(t, t, indexing="ij")
    w = np.linalg.eigvalsh(_green(np.stack([xx.ravel(), yy.ravel()], 1)))
    assert (w < 0).sum() == 3


def test_lu_library_avoids_magma_off_the_default_stream():
    """MAGMA's LU runs on its own queues, ordered only with the default stream: any other
    current stream selects hipSOLVER (torch's "cusolver" name), which runs on that stream."""
    from maomi_amd.numerics import lu_library
    assert lu_library(True) == "magma"
    assert lu_library(False) == "cusolver"


def test_point_grid_and_distance_tables():
    from maomi_amd.numerics import grid_tables, point_grid
    rng = np.random.default_rng(3)
    xs, ys = np.array([0.0, 2.5, 9.0]), np.array([1.0, 4.0])
    full = np.stack(np.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
    perm = rng.permutation(len(full))
    gx, gy, order = point_grid(full[perm])
    np.testing.assert_array_equal(gx, xs)
    np.testing.assert_array_equal(gy, ys)
    # order[j * len(ys) + k] is the index of the point at (xs[j], ys[k])
    np.testing.assert_array_equal(full[perm][order], full)
    assert point_grid(full[:-1]) is None                          # incomplete grid
    assert point_grid(np.concatenate([full[:-1], full[:1]])) is None  # a duplicate
    qx, qy = np.array([0.5, 3.0]), np.array([0.0, 1.0, 7.5])
    du, ix, iy = grid_tables((xs, ys, order), qx, qy)
    np.testing.assert_array_equal(du[ix].reshape(2, 3), np.abs(qx[:, None] - xs[None, :]))
    np.testing.assert_array_equal(du[iy].reshape(3, 2), np.abs(qy[:, None] - ys[None, :]))
    assert np.all(np.diff(du) > 0) and ix.dtype == np.int32
    # more than 128 point rows: the table path declines
    many = np.stack(np.meshgrid([0.0, 1.0], np.arange(129.0), indexing="ij"), -1).reshape(-1, 2)
    assert grid_tables(point_grid(many), qx, qy) is None
