"""Host logic of the GPU input generation (no GPU needed)."""

import numpy as np
import pytest


def _full_grid_disk(X, Y, radius):
    x = np.arange(-X / 2, X / 2)[:, None]
    y = np.arange(-Y / 2, Y / 2)[None, :]
    disk = x * x + y * y <= radius * radius
    nz = np.argwhere(disk)
    if not len(nz):
        return None
    lo, hi = nz.min(0), nz.max(0) + 1
    return disk[lo[0]:hi[0], lo[1]:hi[1]]


@pytest.mark.parametrize("X,Y", [(40, 40), (21, 19), (144, 150), (9, 64), (532, 532)])
def test_disk_runs_match_full_grid_disk(X, Y):
    from maomi_amd.inputs import _disk_runs
    radii = [0.2, 0.5, 0.70710678118654757, 1.0, 2.5, 4.04, 7.3, 11.999999, 30.5, 90.2, 300.0,
             1e4]
    # radii through grid points (x^2 + y^2 = k exactly, up to rounding of the square root)
    radii += [float(np.sqrt(k)) for k in range(1, 2000, 7)]
    radii += [float(np.sqrt(k + 0.25)) for k in range(0, 400, 3)]  # half-integer grids
    for radius in radii:
        want = _full_grid_disk(X, Y, radius)
        got = _disk_runs(X, Y, radius)
        if want is None:
            assert got is None
            continue
        runs, K0, K1, count = got
        assert (K0, K1) == want.shape and count == want.sum()
        rebuilt = np.zeros((K0, K1), bool)
        for k in range(K0):
            rebuilt[k, runs[2 * k]:runs[2 * k + 1] + 1] = True
        np.testing.assert_array_equal(rebuilt, want)


def test_disk_runs_many_equals_one_at_a_time():
    from maomi_amd.inputs import _disk_runs, _disk_runs_many
    radii = [0.2, 3.0, 0.4, 90.2, 17.5, 1e4, 0.6, 55.0]
    for X, Y in [(532, 532), (21, 19)]:
        many = _disk_runs_many(X, Y, radii)
        for r, got in zip(radii, many):
            one = _disk_runs(X, Y, r)
            assert (got is None) == (one is None)
            if got is not None:
                np.testing.assert_array_equal(got[0], one[0])
                assert got[1:] == one[1:]
