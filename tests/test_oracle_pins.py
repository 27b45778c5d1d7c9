"""Pins four oracle behaviours that the other oracle tests leave open, each with
a check that a plausible mistake fails:

* the linear ("trilinear") Ewald weight (include/uniaxial.h:769-804, 937-941):
  on a scatter volume that is affine in (x, y, z) the linear projection must be
  the affine function evaluated at the continuous z index of the Ewald point
  (swapping f and 1 - f, or measuring f from the wrong plane, fails), and the
  nearest-neighbour projection the function at the rounded index;
* the FULL projection (include/uniaxial.h:821-874): projecting the stored
  scatter volume must give exactly the PARTIAL projection from p~, which
  evaluates the same intensity at the same points (include/uniaxial.h:876-947);
* the rotation mask (include/uniaxial.h:958-992): each pixel's average is the
  mean of its finite rotated values, 0 where none is finite, assembled here
  from single-angle runs (counting NaN samples, dividing by the angle count or
  leaving NaN in the sum fails);
* the 1/nA scale (src/cudaMain.cu:897-909): averaging two or three copies of
  one angle (0:360:360, 0:360:720) returns that angle's pattern (dividing by
  nA - 1 or not dividing fails).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_constants, random_euler


def _q(i, n, d):
    return -np.pi / d + i * (2 * np.pi / d) / (n - 1)


def _continuous_z(dims, d, k, kv):
    """Continuous z index (pos_z - start) / dz of each pixel's Ewald point and
    the pixels the reference marks NaN before any z test (val < 0, last row or
    column).  Independent numpy statement of include/uniaxial.h:903-921."""
    nx, ny, nz = dims
    Y, X = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    qx, qy = _q(X, nx, d), _q(Y, ny, d)
    val = k * k - (k * kv[0] + qx) ** 2 - (k * kv[1] + qy) ** 2
    dead = (val < 0) | (X == nx - 1) | (Y == ny - 1)
    posz = -k * kv[2] + np.sqrt(np.where(val < 0, 0, val))
    dz = (2 * np.pi / d) / (nz - 1)
    return X, Y, (posz + np.pi / d) / dz, dead


AFFINE = (0.7, 0.31, -0.023, 0.017)  # S = a + b z + c x + e y


def _affine_volume(dims):
    nx, ny, nz = dims
    a, b, c, e = AFFINE
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return a + b * Z + c * X + e * Y


KCASES = {"kz": (0.0, 0.0, 1.0), "tilted": tuple(np.array([0.08, -0.05, 1.0]) / np.linalg.norm([0.08, -0.05, 1.0]))}


@pytest.mark.parametrize("kname", sorted(KCASES))
@pytest.mark.parametrize("energy", [284.0, 300.0])
def test_linear_weight_exact_on_affine_volume(kname, energy):
    dims, d = (24, 20, 16), 5.0
    kv = KCASES[kname]
    k = 2 * np.pi / (1239.84197 / energy)
    S = _affine_volume(dims)
    X, Y, t, dead = _continuous_z(dims, d, k, kv)
    a, b, c, e = AFFINE
    nz = dims[2]
    sim = O.Sim(dims, 1, d, [energy], (0, 1, 0), interpolation=1)
    got = O.ewald_full(sim, S, k, kv)
    # linear: valid while the bracketing planes Z0 = trunc(t), Z0 + 1 exist
    lin_ok = ~dead & (t > -1) & (np.trunc(t) + 1 <= nz - 1)
    assert lin_ok.sum() > 0.3 * lin_ok.size
    expect = a + b * t + c * X + e * Y
    assert np.all(np.isnan(got[~lin_ok]))
    np.testing.assert_allclose(got[lin_ok], expect[lin_ok], rtol=1e-11, atol=0)
    # the weight is not symmetric on these pixels: a swapped weight would move them
    frac = t[lin_ok] - np.floor(t[lin_ok])
    assert np.abs(frac - 0.5).max() > 0.3
    # nearest neighbour: the plane round(t)
    sim0 = O.Sim(dims, 1, d, [energy], (0, 1, 0), interpolation=0)
    got0 = O.ewald_full(sim0, S, k, kv)
    r = np.round(t)
    nn_ok = ~dead & (r >= 0) & (r <= nz - 1)
    expect0 = a + b * r + c * X + e * Y
    assert np.all(np.isnan(got0[~nn_ok]))
    np.testing.assert_allclose(got0[nn_ok], expect0[nn_ok], rtol=1e-13, atol=0)


def test_linear_2d_reads_plane_zero():
    """In 2D both interpolations read the single plane (no z test)."""
    dims, d, E = (32, 24, 1), 1.0, 285.0
    k = 2 * np.pi / (1239.84197 / E)
    S = np.random.default_rng(3).uniform(1, 2, size=(1, 24, 32))
    for interp in (0, 1):
        got = O.ewald_full(O.Sim(dims, 1, d, [E], (0, 1, 0), interpolation=interp), S, k)
        Y, X = np.meshgrid(np.arange(24), np.arange(32), indexing="ij")
        val = k * k - _q(X, 32, d) ** 2 - _q(Y, 24, d) ** 2
        ok = (val >= 0) & (X != 31) & (Y != 23)
        assert np.array_equal(np.isfinite(got), ok)
        assert np.array_equal(got[ok], S[0][ok])


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("dims", [(20, 18, 12), (24, 24, 1)])
@pytest.mark.parametrize("kname", sorted(KCASES))
def test_full_projection_equals_partial(interp, dims, kname):
    kv = KCASES[kname]
    d = 5.0 if dims[2] > 1 else 1.0
    E = 287.0
    k = 2 * np.pi / (1239.84197 / E)
    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    P = [rng.normal(size=(nz, ny, nx)) + 1j * rng.normal(size=(nz, ny, nx)) for _ in range(3)]
    sim = O.Sim(dims, 1, d, [E], (0, 1, 0), interpolation=interp)
    part = O.ewald_partial(sim, *P, k, kv)
    full = O.ewald_full(sim, O.scatter3d(sim, *P, k, kv), k, kv)
    assert np.isfinite(part).sum() > 0
    assert np.array_equal(part, full, equal_nan=True)


@pytest.mark.parametrize("interp,alg", [(0, 0), (1, 0), (1, 1)])
def test_full_approach_simulation_equals_partial(interp, alg):
    dims, M = (16, 16, 8), 2
    vox = random_euler(dims, M, 5)
    mats = random_constants(2, M, 6)
    out = [O.simulate(O.Sim(dims, M, 5.0, [284.0, 286.0], (0, 40, 120), interpolation=interp, algorithm=alg,
                            scatter_approach=s), vox, mats) for s in (0, 1)]
    assert np.array_equal(out[0], out[1], equal_nan=True)


def _single(dims, M, vox, mats, deg, **kw):
    return O.simulate(O.Sim(dims, M, 5.0, [285.0], (deg, 1.0, deg), **kw), vox, mats)[0, 0]


def test_rotation_mask_single_angle_is_nan_to_zero():
    dims, M = (24, 24, 8), 2
    vox = random_euler(dims, M, 12)
    mats = random_constants(1, M, 13)
    plain = _single(dims, M, vox, mats, 0.0, interpolation=1)
    masked = _single(dims, M, vox, mats, 0.0, interpolation=1, rot_mask=1)
    assert np.isnan(plain).any() and np.isfinite(plain).any()
    assert np.array_equal(masked, np.where(np.isnan(plain), 0.0, plain))


@pytest.mark.parametrize("angles", [(0.0, 180.0, 180.0), (10.0, 35.0, 80.0), (0.0, 90.0, 270.0)])
def test_rotation_mask_average_counts_finite_samples(angles):
    """Mask average over several angles = per-pixel mean of the finite single-angle
    results (each single-angle run is one rotated projection, unscaled)."""
    dims, M = (24, 24, 8), 2
    vox = random_euler(dims, M, 14)
    mats = random_constants(1, M, 15)
    n = O.num_angles(*angles)
    singles = np.stack([_single(dims, M, vox, mats, angles[0] + i * angles[1], interpolation=1) for i in range(n)])
    fin = np.isfinite(singles)
    count = fin.sum(axis=0)
    total = np.where(fin, singles, 0.0).sum(axis=0)
    expect = np.where(count == 0, 0.0, total / np.maximum(count, 1))
    got = O.simulate(O.Sim(dims, M, 5.0, [285.0], angles, interpolation=1, rot_mask=1), vox, mats)[0, 0]
    # pixels with a partial count exist (so dividing by n instead of the count would fail)
    assert np.any((count > 0) & (count < n))
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(got, expect, rtol=1e-13, atol=0)
    # without the mask the average is NaN wherever any sample is NaN, else the mean over n
    nomask = O.simulate(O.Sim(dims, M, 5.0, [285.0], angles, interpolation=1), vox, mats)[0, 0]
    all_fin = count == n
    assert np.array_equal(np.isfinite(nomask), all_fin)
    np.testing.assert_allclose(nomask[all_fin], total[all_fin] / n, rtol=1e-13, atol=0)


@pytest.mark.parametrize("angles,n", [((0.0, 360.0, 360.0), 2), ((0.0, 360.0, 720.0), 3)])
def test_average_of_repeated_angle_is_that_angle(angles, n):
    dims, M = (24, 24, 8), 2
    vox = random_euler(dims, M, 16)
    mats = random_constants(1, M, 17)
    assert O.num_angles(*angles) == n
    one = _single(dims, M, vox, mats, 0.0, interpolation=1)
    avg = O.simulate(O.Sim(dims, M, 5.0, [285.0], angles, interpolation=1), vox, mats)[0, 0]
    # 360 degrees is 0 up to the rounding of cos/sin(2 pi): tiny back-map offsets can
    # drop border pixels, so compare where both are finite
    both = np.isfinite(one) & np.isfinite(avg)
    assert both.sum() > 0.8 * np.isfinite(one).sum()
    np.testing.assert_allclose(avg[both], one[both], rtol=1e-9, atol=0)
