"""Pins the oracle's rotation math, polarization, Ewald geometry, warp and
whole-pipeline behaviour with closed forms, invariants and physics.

None of these restate the oracle's code: each checks a property the mathematics
fixes (a rotation is orthonormal and maps x to the E direction, a uniaxial
tensor reduces to n^2 - 1 for aligned/isotropic cases, the Ewald sample lies on
the sphere |k k^ + q| = k, bilinear interpolation reproduces affine images, the
Born intensity scales as |n^2 - 1|^2, ...).
"""
import numpy as np
import pytest

from cyrsoxs_amd import synthetic
from oracle import oracle as O


def Rz(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


# ---------------------------------------------------------------- rotations
@pytest.mark.parametrize("ang", [0.0, 0.3, 1.0, np.pi / 2, 2.5, np.pi, 4.0, 2 * np.pi - 0.01])
def test_e_rotation_is_rz_for_default_case(ang):
    R = O.rotation_matrix_e((0, 0, 1), ang)
    assert np.abs(R - Rz(ang)).max() < 1e-12


@pytest.mark.parametrize("seed", range(6))
def test_vec_to_vec_rotation(seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=3)
    b = rng.normal(size=3)
    a /= np.linalg.norm(a)
    b /= np.linalg.norm(b)
    R = O.rotation_vec_to_vec(a, b)
    assert np.allclose(R @ a, b, atol=1e-12)
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-12)
    assert np.isclose(np.linalg.det(R), 1.0)
    # rotation axis is a x b: it is left fixed
    ax = np.cross(a, b)
    assert np.allclose(R @ ax, ax, atol=1e-12)
    assert np.array_equal(O.rotation_vec_to_vec(a, a), np.eye(3))  # FEQUALS shortcut


def test_rodrigues_and_base_configuration():
    k = np.array([0.0, 0.5, 0.5])
    k /= np.linalg.norm(k)
    v = np.array([1.0, 0, 0])
    r = O.rodrigues(v, k, 0.7)
    assert np.isclose(np.dot(r, k), np.dot(v, k))
    assert np.isclose(np.linalg.norm(r), 1.0)
    RK = O.rotation_matrix_k(k)
    assert np.allclose(RK @ [0, 0, 1], k)
    R, ang = O.base_configuration(k)
    x = R @ [1, 0, 0]
    assert abs(x[1]) < 1e-4 and x[0] > 0  # rotated x has |y| minimal and x > 0
    assert np.allclose(R @ [0, 0, 1], k, atol=1e-4)
    for t in (0.2, 1.3):  # E = R(t) x lies in the plane normal to k
        E = O.rotation_matrix_e(k, t) @ [1, 0, 0]
        assert abs(np.dot(E, k)) < 1e-12


def test_warp_affine_matrix_recovers_affine():
    A = np.array([[0.8, -0.3, 5.0], [0.2, 1.1, -2.0]])
    src = np.array([[1.0, 2.0], [7.0, 3.0], [4.0, 9.0]])
    dst = (A[:, :2] @ src.T).T + A[:, 2]
    c = O.warp_affine_matrix(src, dst)
    assert np.allclose(c.reshape(2, 3), A, atol=1e-12)


# ------------------------------------------------------------- polarization
def _one_voxel_sim(morph=0, frame=1):
    return O.Sim((2, 2, 2), 1, 5.0, [285.0], (0, 1, 0), morphology_type=morph, reference_frame=frame)


def _pol(sim, vox, mat, R):
    V = 8
    v = np.repeat(np.asarray(vox, float)[None, None, :], V, axis=1)
    return O.polarization(sim, v, np.asarray(mat, float)[None, :], R)


NPAR = 1 - 2e-3 + 1e-3j
NPER = 1 - 1e-3 + 3e-4j


def _mat(npar=NPAR, nper=NPER):
    return [npar.real, npar.imag, nper.real, nper.imag]


def test_polarization_vacuum_is_zero():
    p = _pol(_one_voxel_sim(), [0.7, 1.1, 0.4, 0.9], _mat(1 + 0j, 1 + 0j), np.eye(3))
    assert max(np.abs(x).max() for x in p) < 1e-15


def test_polarization_isotropic():
    vf = 0.6
    p = _pol(_one_voxel_sim(), [0.0, 0.0, 0.0, vf], _mat(), Rz(0.4))
    nbar = (NPAR + 2 * NPER) / 3
    E = Rz(0.4) @ [1, 0, 0]
    expect = vf * (nbar ** 2 - 1) * E / (4 * np.pi)
    for c in range(3):
        assert np.allclose(p[c], expect[c], rtol=1e-13, atol=1e-18)


def test_polarization_single_aligned_voxel():
    # director along x (theta = pi/2, psi = 0), fully aligned, E along x
    p = _pol(_one_voxel_sim(), [1.0, np.pi / 2, 0.0, 1.0], _mat(), np.eye(3))
    assert np.allclose(p[0], (NPAR ** 2 - 1) / (4 * np.pi), rtol=1e-12)
    assert np.abs(p[1]).max() < 1e-18 and np.abs(p[2]).max() < 1e-18
    # director along z: E along x sees n_perp
    p = _pol(_one_voxel_sim(), [1.0, 0.0, 0.0, 1.0], _mat(), np.eye(3))
    assert np.allclose(p[0], (NPER ** 2 - 1) / (4 * np.pi), rtol=1e-12)


def test_polarization_vector_morphology_closed_form():
    s = np.array([0.5, -0.2, 0.3])
    pu = 0.25
    p = _pol(_one_voxel_sim(morph=1), [*s, pu], _mat(), np.eye(3))
    # tensor: n_perp^2 I |s|^2 + (n_par^2 - n_perp^2) s s^T + pu nbar^2 I - (pu + |s|^2) I
    nbar2 = ((NPAR + 2 * NPER) / 3) ** 2
    N = (NPER ** 2) * np.eye(3) * (s @ s) + (NPAR ** 2 - NPER ** 2) * np.outer(s, s) + pu * nbar2 * np.eye(3) \
        - (pu + s @ s) * np.eye(3)
    expect = N @ [1, 0, 0] / (4 * np.pi)
    for c in range(3):
        assert np.allclose(p[c], expect[c], rtol=1e-12, atol=1e-18)


def test_polarization_linear_in_e_and_material_frame():
    rng = np.random.default_rng(3)
    vox = [0.8, 1.0, 2.0, 0.7]
    sim = _one_voxel_sim()
    p0 = np.array([x[0] for x in _pol(sim, vox, _mat(), Rz(0.0))])
    p90 = np.array([x[0] for x in _pol(sim, vox, _mat(), Rz(np.pi / 2))])
    t = rng.uniform(0, 2 * np.pi)
    pt = np.array([x[0] for x in _pol(sim, vox, _mat(), Rz(t))])
    assert np.allclose(pt, np.cos(t) * p0 + np.sin(t) * p90, rtol=1e-12, atol=1e-18)
    pm = np.array([x[0] for x in _pol(_one_voxel_sim(frame=0), vox, _mat(), Rz(t))])
    assert np.allclose(pm, Rz(t).T @ pt, rtol=1e-12, atol=1e-18)


def test_nt_path_equals_direct_polarization():
    dims = (6, 5, 4)
    rng = np.random.default_rng(5)
    M = 3
    vox = np.zeros((M, 120, 4))
    vox[:, :, 0] = rng.uniform(0, 1, (M, 120))
    vox[:, :, 1] = rng.uniform(0, np.pi, (M, 120))
    vox[:, :, 2] = rng.uniform(0, 2 * np.pi, (M, 120))
    vox[:, :, 3] = rng.dirichlet(np.ones(M), 120).T
    mats = np.array([_mat(1 - 1e-3 * (m + 1) + 1e-4j, 1 - 5e-4 + 2e-4j) for m in range(M)])
    for frame in (0, 1):
        sim = O.Sim(dims, M, 5.0, [285.0], (0, 1, 0), reference_frame=frame)
        R = Rz(0.9)
        a = O.polarization(sim, vox, mats, R)
        b = O.polarization_nt(sim, O.compute_nt(sim, vox, mats), R)
        for x, y in zip(a, b):
            assert np.allclose(x, y, rtol=1e-12, atol=1e-18)


def test_hanning_window():
    dims = (8, 6, 4)
    vox = np.zeros((1, 8 * 6 * 4, 4))
    vox[0, :, 3] = 1.0
    s0 = O.Sim(dims, 1, 5.0, [285.0], (0, 1, 0), windowing=0)
    s1 = O.Sim(dims, 1, 5.0, [285.0], (0, 1, 0), windowing=1)
    m = np.array([_mat()])
    a = O.polarization(s0, vox, m, np.eye(3))[0].reshape(4, 6, 8)
    b = O.polarization(s1, vox, m, np.eye(3))[0].reshape(4, 6, 8)
    h = lambda n: 0.5 * (1 - np.cos(2 * np.pi * np.arange(n) / n))
    w = h(4)[:, None, None] * h(6)[None, :, None] * h(8)[None, None, :]
    assert np.allclose(b, a * w, rtol=1e-13, atol=1e-20)


# ------------------------------------------------------------- Ewald geometry
def _plane_volume(dims, plane_z):
    """p~ = (1, 0, 0) on one z plane of the SHIFTED volume, 0 elsewhere."""
    nx, ny, nz = dims
    S = np.zeros((nz, ny, nx), complex)
    S[plane_z] = 1.0
    return S


@pytest.mark.parametrize("interp", [0, 1])
def test_ewald_samples_lie_on_sphere(interp):
    dims, d, E = (32, 32, 32), 5.0, 285.0
    nx, ny, nz = dims
    sim = O.Sim(dims, 1, d, [E], (0, 1, 0), interpolation=interp)
    k = 2 * np.pi / (1239.84197 / E)
    q = lambda i, n: -np.pi / d + i * (2 * np.pi / d) / (n - 1)
    zero = np.zeros((nz, ny, nx), complex)
    hits = np.zeros((ny, nx), int)
    for z in range(nz):
        proj = O.ewald_partial(sim, _plane_volume(dims, z), zero, zero, k)
        lit = np.isfinite(proj) & (proj > 0)
        hits += lit
        Y, X = np.nonzero(lit)
        qz = -k + np.sqrt(k * k - q(X, nx) ** 2 - q(Y, ny) ** 2)  # the sphere |k z^ + q| = k
        dz = (2 * np.pi / d) / (nz - 1)
        # the plane sampled is the nearest grid plane (NN) or a bracketing one (linear)
        assert np.all(np.abs(qz - q(z, nz)) <= (0.5 if interp == 0 else 1.0) * dz * (1 + 1e-12))
    valid = np.isfinite(O.ewald_partial(sim, zero + 1, zero, zero, k))
    assert np.all(hits[valid] == (1 if interp == 0 else 2) - (hits[valid] == 1) * (interp == 1) * 0 + 0) or interp
    # NaN structure: last row and column, and points outside the sphere's footprint
    assert np.all(~valid[-1, :]) and np.all(~valid[:, -1])
    Yg, Xg = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    outside = k * k - q(Xg, nx) ** 2 - q(Yg, ny) ** 2 < 0
    assert np.all(~valid[outside])


def test_ewald_intensity_is_transverse_projection():
    dims, d, E = (16, 16, 16), 5.0, 290.0
    nx, ny, nz = dims
    rng = np.random.default_rng(7)
    P = [rng.normal(size=(nz, ny, nx)) + 1j * rng.normal(size=(nz, ny, nx)) for _ in range(3)]
    sim = O.Sim(dims, 1, d, [E], (0, 1, 0), interpolation=0)
    k = 2 * np.pi / (1239.84197 / E)
    S = O.scatter3d(sim, *P, k).reshape(nz, ny, nx)
    q = lambda i, n: -np.pi / d + i * (2 * np.pi / d) / (n - 1)
    for (z, y, x) in [(3, 4, 5), (8, 8, 8), (0, 15, 2)]:
        v = np.array([q(x, nx), q(y, ny), k + q(z, nz)])
        p = np.array([P[0][z, y, x], P[1][z, y, x], P[2][z, y, x]])
        expect = np.linalg.norm((k * k * np.eye(3) - np.outer(v, v)) @ p) ** 2
        assert np.isclose(S[z, y, x], expect, rtol=1e-12)


# ------------------------------------------------------------------- warp
def test_warp_identity_and_quarter_turns():
    rng = np.random.default_rng(9)
    n = 32
    img = rng.uniform(size=(n, n))
    nan_img = img.copy()
    nan_img[3, 4] = np.nan
    # 0 degrees: cos = 1, sin = 0 exactly, every tap but one has weight 0 -> exact copy
    assert np.array_equal(O.warp_affine(nan_img, O.rotation_coeffs(0.0, n, n)), nan_img, equal_nan=True)
    # 90 degrees about (n/2, n/2): on integer pixels a permutation (up to the
    # 1e-16 of cos(pi/2) in double); destination pixels mapping outside stay NaN
    r90 = O.warp_affine(img, O.rotation_coeffs(np.pi / 2, n, n))
    c = n // 2
    Y, X = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    sx, sy = 2 * c - Y, X
    inner = (sx >= 1) & (sx <= n - 2) & (sy >= 1) & (sy <= n - 2)
    assert np.allclose(r90[inner], img[sy[inner], sx[inner]], rtol=0, atol=1e-12)
    outer = (sx < 0) | (sx > n - 1) | (sy < 0) | (sy > n - 1)
    assert np.all(np.isnan(r90[outer]))


@pytest.mark.parametrize("deg", [13.0, 47.0, 150.0])
def test_warp_reproduces_affine_images(deg):
    n = 40
    Y, X = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing="ij")
    img = 0.3 * X - 1.7 * Y + 2.0
    t = np.deg2rad(deg)
    out = O.warp_affine(img, O.rotation_coeffs(t, n, n))
    c = n / 2
    sx = c + np.cos(t) * (X - c) - np.sin(t) * (Y - c)
    sy = c + np.sin(t) * (X - c) + np.cos(t) * (Y - c)
    inside = (sx >= 1e-9) & (sx <= n - 1 - 1e-9) & (sy >= 1e-9) & (sy <= n - 1 - 1e-9)
    expect = 0.3 * sx - 1.7 * sy + 2.0
    assert np.allclose(out[inside], expect[inside], atol=1e-9)
    outside = (sx < -1e-9) | (sx > n - 1 + 1e-9) | (sy < -1e-9) | (sy > n - 1 + 1e-9)
    assert np.all(np.isnan(out[outside]))


# --------------------------------------------------------------- whole pipeline
def _iso_two_phase(n=32, r=6.0):
    c = n / 2
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inside = (np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) <= r).astype(float)
    vox = np.zeros((2, n ** 3, 4))
    vox[0, :, 3] = inside.ravel()
    vox[1, :, 3] = 1 - inside.ravel()
    return vox


def test_contrast_scaling():
    """Born intensity is quadratic in the contrast: (n1^2 - 1) -> a (n1^2 - 1) with
    material 2 = vacuum scales every pattern by |a|^2."""
    n = 32
    vox = _iso_two_phase(n)
    sim = O.Sim((n, n, n), 2, 5.0, [285.0], (0, 45, 90), interpolation=1)
    n1 = 1 - 1e-3 + 4e-4j
    a = 1.7 - 0.6j
    n1b = np.sqrt(1 + a * (n1 ** 2 - 1))
    m1 = np.array([[[n1.real, n1.imag, n1.real, n1.imag], [1, 0, 1, 0]]])
    m2 = np.array([[[n1b.real, n1b.imag, n1b.real, n1b.imag], [1, 0, 1, 0]]])
    I1 = O.simulate(sim, vox, m1)
    I2 = O.simulate(sim, vox, m2)
    f = np.isfinite(I1)
    assert np.array_equal(f, np.isfinite(I2))
    assert np.allclose(I2[f], abs(a) ** 2 * I1[f], rtol=1e-9)


def test_sphere_form_factor():
    """An isotropic sphere's pattern follows the Born form factor
    [3 (sin qR - qR cos qR) / (qR)^3]^2 inside the first lobe.  q is taken from the
    FFT bin each detector pixel reads (|q| = 2 pi |kbin| / (N d))."""
    n, r, d, E = 64, 8.0, 5.0, 285.0
    vox = _iso_two_phase(n, r)
    nbar = 1 - 1e-3 + 3e-4j
    mats = np.array([[[nbar.real, nbar.imag, nbar.real, nbar.imag], [1, 0, 1, 0]]])
    sim = O.Sim((n, n, n), 2, d, [E], (0, 1, 0), interpolation=0)
    I = O.simulate(sim, vox, mats)[0, 0]
    k = 2 * np.pi / (1239.84197 / E)
    # pixels along the row through the centre (Y = n/2): bin ky = 0, kx = n/2 - X
    Y = n // 2
    R = r * d
    vals, ffs = [], []
    for X in range(n // 2 - 5, n // 2):
        kb = n // 2 - X
        qq = 2 * np.pi * kb / (n * d)
        qx_lab = -np.pi / d + X * (2 * np.pi / d) / (n - 1)
        pol = k ** 4 - k * k * qx_lab ** 2  # |(k^2 I - v v^T) x^|^2 to leading order, E = x^
        ff = (3 * (np.sin(qq * R) - qq * R * np.cos(qq * R)) / (qq * R) ** 3) ** 2
        vals.append(I[Y, X] / pol)
        ffs.append(ff)
    vals = np.array(vals) / vals[-1]
    ffs = np.array(ffs) / ffs[-1]
    assert np.allclose(vals, ffs, rtol=0.05, atol=3e-3), (vals, ffs)


def test_sphere_pattern_symmetry():
    """A radial-director sphere averaged over 0:30:150 is symmetric under the
    detector's 180-degree rotation about the centre (pairs of angles differ by 180)."""
    c = synthetic.config0()
    sim = O.Sim(c["dims"], 2, c["phys_size"], [285.0], (0.0, 90.0, 90.0), interpolation=0)
    I = O.simulate(sim, c["voxels"], c["mats"][1:2])[0, 0]
    assert np.isfinite(I[32, 32]) and I[32, 32] > 0
    assert np.all(np.isnan(I[-1, :])) and np.all(np.isnan(I[:, -1]))
