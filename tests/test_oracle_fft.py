"""Pins the oracle's FFT, DC replacement and Igor shift.

FFT: against numpy.fft (a library routine), the brute-force O(N^2) DFT, and
closed forms (delta -> constant, plane wave -> one bin, Parseval).  Shift and
DC: against the integer index maps stated independently with numpy indexing.
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", list(range(1, 41)) + [48, 64, 96, 97, 128, 243, 384, 625, 1024, 2048])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    ref = np.fft.fft(x)
    got = O.fft1d(x)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 8, 12, 16, 30, 64])
def test_fft1d_matches_direct_dft(n):
    rng = np.random.default_rng(100 + n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.allclose(O.fft1d(x), O.dft1d_direct(x), atol=1e-12)
    # the direct DFT itself against the definition written as a matrix product
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.allclose(O.dft1d_direct(x), W @ x, atol=1e-11)


@pytest.mark.parametrize("shape", [(4, 6, 8), (16, 16, 16), (3, 5, 7), (1, 32, 32), (12, 10, 48), (2, 3, 384)])
def test_fft3d_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    ref = np.fft.fftn(v)
    got = O.fft3d(v)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_fft_closed_forms():
    nz, ny, nx = 8, 12, 16
    d = np.zeros((nz, ny, nx), complex)
    d[0, 0, 0] = 2.5
    assert np.allclose(O.fft3d(d), 2.5)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky, kz = 3, 5, 1
    pw = np.exp(2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
    F = O.fft3d(pw)
    expect = np.zeros_like(F)
    expect[kz, ky, kx] = nx * ny * nz
    assert np.abs(F - expect).max() < 1e-9
    rng = np.random.default_rng(0)
    v = rng.normal(size=(nz, ny, nx)) + 1j * rng.normal(size=(nz, ny, nx))
    assert np.isclose(np.sum(np.abs(O.fft3d(v)) ** 2), v.size * np.sum(np.abs(v) ** 2), rtol=1e-12)


@pytest.mark.parametrize("shape", [(8, 8, 8), (5, 6, 7), (1, 16, 16), (4, 9, 10)])
def test_igor_shift_index_map(shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(1)
    F = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    expect = F[(nz // 2 - Z) % nz, (ny // 2 - Y) % ny, (nx // 2 - X) % nx]
    got = O.igor_shift(F)
    assert np.array_equal(got, expect)
    assert np.array_equal(O.igor_shift(got), F)  # involution
    # where DC lands: F[0,0,0] -> (nx/2, ny/2, nz/2)
    assert got[nz // 2, ny // 2, nx // 2] == F[0, 0, 0]


@pytest.mark.parametrize("shape", [(8, 8, 8), (1, 16, 16), (3, 4, 5)])
def test_dc_replacement_rule(shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(2)
    F = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    got = O.dc_replace(F)
    nb = [F[0, 0, 1], F[0, 1, 0], F[0, 0, nx - 1], F[0, ny - 1, 0]]
    if nz > 1:
        nb += [F[1, 0, 0], F[nz - 1, 0, 0]]
    assert np.isclose(got[0, 0, 0], np.mean(nb), rtol=1e-15, atol=0)
    got[0, 0, 0] = F[0, 0, 0]
    assert np.array_equal(got, F)
