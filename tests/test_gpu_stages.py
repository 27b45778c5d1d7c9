"""Per-stage parity of the HIP kernels against the FP64 oracle (through the C ABI).

Tolerances: FP64 <= 1e-12 relative L2 for single stages, FP32 <= 1e-5 (FFT,
polarization) and <= 1e-5 (Ewald, warp); NaN masks and nearest-neighbour bins
must match exactly.
"""
import numpy as np
import pytest

from cyrsoxs_amd import capi
from oracle import oracle as O
from tests.helpers import random_constants, random_euler, random_vector, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FFT_SHAPES = [(64, 64, 64), (16, 12, 10), (384, 8, 3), (48, 48, 48), (128, 128, 32), (256, 256, 1),
              (2048, 16, 1), (7, 11, 13), (30, 20, 5), (512, 4, 2),
              # compile-time codelets on strided axes with fewer pencils than a tile
              (8, 8, 512), (6, 1024, 2), (5, 384, 3)]


def _cplx(prec):
    return (torch.complex128, np.complex128) if prec == capi.FP64 else (torch.complex64, np.complex64)


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("dims", FFT_SHAPES)
def test_fft3d_matches_oracle(dims, prec):
    nx, ny, nz = dims
    print(f"fft3d dims={dims} prec={prec}", flush=True)
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    vols = [rng.normal(size=(nz, ny, nx)) + 1j * rng.normal(size=(nz, ny, nx)) for _ in range(3)]
    tdt, ndt = _cplx(prec)
    dev = [torch.from_numpy(v.astype(ndt)).cuda() for v in vols]
    capi.fft3d_device(dev, dims, prec)
    torch.cuda.synchronize()
    tol = 1e-12 if prec == capi.FP64 else 2e-6
    for v, d in zip(vols, dev):
        ref = O.fft3d(v)
        err = rel_l2(d.cpu().numpy(), ref)
        assert err < tol, (dims, prec, err)


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("dims", [(16, 12, 10), (64, 64, 1), (8, 8, 8)])
def test_dc_and_shift(dims, prec):
    nx, ny, nz = dims
    rng = np.random.default_rng(5)
    v = rng.normal(size=(nz, ny, nx)) + 1j * rng.normal(size=(nz, ny, nx))
    tdt, ndt = _cplx(prec)
    d = torch.from_numpy(v.astype(ndt)).cuda()
    capi.dc_shift_device(d, dims, prec, True, True)
    ref = O.igor_shift(O.dc_replace(v))
    got = d.cpu().numpy().astype(np.complex128)
    tol = 1e-14 if prec == capi.FP64 else 1e-6
    assert rel_l2(got, ref) < tol


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("morph,frame,window", [(0, 1, 0), (0, 0, 0), (0, 1, 1), (1, 1, 0), (1, 0, 1)])
def test_polarization_matches_oracle(morph, frame, window, prec):
    dims = (12, 10, 6)
    M = 3
    vox = random_euler(dims, M, 11) if morph == 0 else random_vector(dims, M, 11)
    mats = random_constants(2, M, 3)
    kw = dict(morphology_type=morph, reference_frame=frame, windowing=window)
    sim = capi.Sim(dims, M, 5.0, [280.0, 281.0], (0, 30, 150), precision=prec, **kw)
    osim = O.Sim(dims, M, 5.0, [280.0, 281.0], (0, 30, 150), **kw)
    for ang in (0.0, 0.7, np.pi):
        px, py, pz = capi.polarization(sim, vox, mats, 1, ang)
        R = O.rotation_matrix_e((0, 0, 1), ang)
        rx, ry, rz = O.polarization(osim, vox, mats[1], R)
        tol = 1e-12 if prec == capi.FP64 else 2e-5
        for g, r in ((px, rx), (py, ry), (pz, rz)):
            assert rel_l2(g, r) < tol, (morph, frame, window, prec, ang, rel_l2(g, r))


def _fft_volumes(dims, seed):
    """Random polarization-like volumes after FFT + DC replacement (unshifted)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        v = (rng.normal(size=(nz, ny, nx)) + 1j * rng.normal(size=(nz, ny, nx))) * 1e-3
        out.append(O.dc_replace(O.fft3d(v)))
    return out


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("dims,phys,interp,full", [((32, 32, 16), 5.0, 0, 0), ((32, 32, 16), 5.0, 1, 0),
                                                   ((32, 32, 16), 5.0, 1, 1), ((40, 40, 1), 1.0, 1, 0),
                                                   ((24, 24, 24), 2.0, 0, 1), ((64, 64, 64), 5.0, 0, 0)])
def test_ewald_matches_oracle(dims, phys, interp, full, prec):
    nx, ny, nz = dims
    F = _fft_volumes(dims, 17)
    energies = [280.0, 290.0]
    sim = capi.Sim(dims, 1, phys, energies, (0, 1, 0), interpolation=interp, scatter_approach=full, precision=prec)
    osim = O.Sim(dims, 1, phys, energies, (0, 1, 0), interpolation=interp, scatter_approach=full)
    tdt, ndt = _cplx(prec)
    dev = [torch.from_numpy(f.astype(ndt)).cuda() for f in F]
    rdt = torch.float64 if prec == capi.FP64 else torch.float32
    for e, E in enumerate(energies):
        proj = torch.zeros((ny, nx), dtype=rdt, device="cuda")
        capi.ewald_device(sim, e, 0, *dev, proj)
        k = 2 * np.pi / (1239.84197 / E)
        S = [O.igor_shift(f) for f in F]
        if full:
            ref = O.ewald_full(osim, O.scatter3d(osim, *S, k), k)
        else:
            ref = O.ewald_partial(osim, *S, k)
        got = proj.cpu().numpy().astype(np.float64)
        assert np.array_equal(np.isnan(got), np.isnan(ref)), "NaN mask differs"
        m = np.isfinite(ref)
        tol = 1e-12 if prec == capi.FP64 else 1e-5
        assert rel_l2(got, ref, m) < tol, rel_l2(got, ref, m)


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("n,ang", [(64, 0.0), (64, 90.0), (64, 30.0), (65, 17.0), (32, 150.0), (64, 180.0)])
def test_warp_matches_oracle(n, ang, prec):
    rng = np.random.default_rng(n + int(ang))
    src = rng.uniform(0, 1, (n, n))
    src[rng.uniform(0, 1, (n, n)) < 0.05] = np.nan
    src[-1, :] = np.nan
    src[:, -1] = np.nan
    c = O.rotation_coeffs(np.deg2rad(ang), n, n)
    ref = O.warp_affine(src, c)
    rdt = torch.float64 if prec == capi.FP64 else torch.float32
    s = torch.from_numpy(src).to(rdt).cuda()
    d = torch.full((n, n), float("nan"), dtype=rdt, device="cuda")
    capi.warp_affine_device(s, d, c, prec)
    got = d.cpu().numpy().astype(np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(ref)), "rotation mask differs"
    m = np.isfinite(ref)
    tol = 1e-14 if prec == capi.FP64 else 1e-6
    assert rel_l2(got, ref, m) < tol
