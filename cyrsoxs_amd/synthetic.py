"""Seeded synthetic morphologies and optical constants for the benchmark workloads.

The reference ships no benchmark inputs (its Data/ folder holds one 16x32x32
vector morphology without expected outputs), so workloads are generated here:

  sphere_morphology      2 materials, centred sphere with a radial director plus its
                         isotropic complement (the 64^3 plain-method case)
  bulk_morphology        M materials, Gaussian-filtered white noise -> softmax volume
                         fractions, filtered uniform S, sphere-uniform Euler angles
  thinfilm_morphology    3-phase 2D spinodal-like field with in-plane director tangent
                         to the interfaces
  lorentzian_constants   delta/beta per material from a seeded sum of 3 Lorentzians
  sphere_constants       Lorentzian-scaled anisotropic material + flat isotropic material

Morphology arrays are float32/float64 [M][Nz*Ny*Nx][4] with X fastest, each voxel
holding (S, theta, psi, Vfrac) as the reference's Voxel.s1 for Euler angles
(include/Input/Input.h:61-76).  Optical constants are float64 [nE][M][4] holding
(npara.re, npara.im, nperp.re, nperp.im) with n = (1 - delta) + i*beta
(include/Input/InputData.h:371-375).

Fields are generated with torch so the same code runs on the CPU (tests) and on
the GPU (the benchmark's 512x512x128 volumes).
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _gen(seed: int, *keys: int) -> torch.Generator:
    g = torch.Generator()
    h = seed & 0xFFFFFFFF
    for k in keys:
        h = (h * 1000003 + (k & 0xFFFFFFFF) + 0x9E3779B9) & 0xFFFFFFFFFFFF
    g.manual_seed(h)
    return g


def _randn(shape, seed, keys, device):
    g = _gen(seed, *keys)
    return torch.randn(shape, generator=g, dtype=torch.float32).to(device)


def _rand(shape, seed, keys, device):
    g = _gen(seed, *keys)
    return torch.rand(shape, generator=g, dtype=torch.float32).to(device)


def gaussian_filter_periodic(field: torch.Tensor, sigma: float) -> torch.Tensor:
    """Periodic Gaussian filter of a (Nz,Ny,Nx) or (Ny,Nx) field via the FFT."""
    dims = tuple(range(-field.dim(), 0))
    spec = torch.fft.rfftn(field, dim=dims)
    shape = field.shape
    mult = torch.ones(spec.shape, dtype=torch.float32, device=field.device)
    for ax, n in enumerate(shape):
        if ax == len(shape) - 1:
            f = torch.fft.rfftfreq(n, device=field.device)
        else:
            f = torch.fft.fftfreq(n, device=field.device)
        g = torch.exp(-2.0 * (math.pi * sigma * f) ** 2)
        view = [1] * len(shape)
        view[ax] = -1
        mult = mult * g.view(view)
    return torch.fft.irfftn(spec * mult, s=shape, dim=dims)


def _standardise(x):
    return (x - x.mean()) / (x.std() + 1e-12)


def bulk_morphology(dims, num_materials, seed=1, sigma=6.0, device="cpu", dtype=torch.float32):
    """Gaussian-filtered Vfrac (softmax, sums to 1), filtered uniform S rescaled to
    [0,1], Euler angles uniform on the sphere.  Returns torch [M][V][4]."""
    nx, ny, nz = dims
    shape = (nz, ny, nx)
    V = nx * ny * nz
    logits = torch.stack([
        _standardise(gaussian_filter_periodic(_randn(shape, seed, (m, 0), device), sigma))
        for m in range(num_materials)])
    vfrac = torch.softmax(logits, dim=0).reshape(num_materials, V)
    del logits
    out = torch.empty((num_materials, V, 4), dtype=dtype, device=device)
    for m in range(num_materials):
        s = gaussian_filter_periodic(_rand(shape, seed, (m, 1), device), sigma).reshape(V)
        s = (s - s.min()) / (s.max() - s.min() + 1e-12)
        u = _rand((V,), seed, (m, 2), device)
        v = _rand((V,), seed, (m, 3), device)
        theta = torch.arccos(1 - 2 * u)
        psi = 2 * math.pi * v
        out[m, :, 0] = s.to(dtype)
        out[m, :, 1] = theta.to(dtype)
        out[m, :, 2] = psi.to(dtype)
        out[m, :, 3] = vfrac[m].to(dtype)
    # The reference's readers zero the angles where S == 0
    # (include/PyClass/VoxelData.h:154-160, include/Input/readH5.h:486-519).
    zero = out[:, :, 0] == 0
    out[:, :, 1][zero] = 0
    out[:, :, 2][zero] = 0
    return out


def sphere_morphology(n=64, radius=20.0, dtype=np.float64):
    """Material 1: sphere of the given radius centred at (n/2, n/2, n/2), Vfrac 1
    inside, S = 1, radial director; material 2: its isotropic complement."""
    c = n / 2.0
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rx, ry, rz = x - c, y - c, z - c
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    inside = (r <= radius).astype(np.float64)
    rs = np.where(r > 0, r, 1.0)
    theta = np.where(r > 0, np.arccos(np.clip(rz / rs, -1, 1)), 0.0)
    psi = np.where(r > 0, np.arctan2(ry, rx), 0.0)
    V = n ** 3
    out = np.zeros((2, V, 4), dtype=dtype)
    out[0, :, 0] = 1.0
    out[0, :, 1] = theta.reshape(V)
    out[0, :, 2] = psi.reshape(V)
    out[0, :, 3] = inside.reshape(V)
    out[1, :, 3] = (1.0 - inside).reshape(V)
    return out


def sphere_constants(energies):
    """Material 1: delta_par=2e-3, beta_par=1e-3, delta_perp=1e-3, beta_perp=3e-4
    scaled by L(E) = 1/(1+((E-285)/0.5)^2); material 2: delta=beta=5e-4 flat."""
    e = np.asarray(energies, dtype=np.float64)
    out = np.zeros((len(e), 2, 4))
    L = 1.0 / (1.0 + ((e - 285.0) / 0.5) ** 2)
    dpar, bpar, dper, bper = 2e-3 * L, 1e-3 * L, 1e-3 * L, 3e-4 * L
    out[:, 0, 0] = 1 - dpar
    out[:, 0, 1] = bpar
    out[:, 0, 2] = 1 - dper
    out[:, 0, 3] = bper
    out[:, 1, 0] = 1 - 5e-4
    out[:, 1, 1] = 5e-4
    out[:, 1, 2] = 1 - 5e-4
    out[:, 1, 3] = 5e-4
    return out


def lorentzian_constants(energies, num_materials, seed=1, lo=1e-4, hi=3e-3):
    """(delta_par, beta_par, delta_perp, beta_perp) per material: seeded sum of 3
    Lorentzians over the energy list, affine-mapped into [lo, hi]."""
    e = np.asarray(energies, dtype=np.float64)
    rng = np.random.default_rng(seed + 7919)
    out = np.zeros((len(e), num_materials, 4))
    e0, e1 = float(e.min()), float(e.max())
    for m in range(num_materials):
        for c in range(4):
            f = np.zeros_like(e)
            for _ in range(3):
                centre = rng.uniform(e0, e1 if e1 > e0 else e0 + 1)
                width = rng.uniform(0.5, 3.0)
                amp = rng.uniform(0.2, 1.0)
                f += amp / (1.0 + ((e - centre) / width) ** 2)
            if f.max() > f.min():
                f = (f - f.min()) / (f.max() - f.min())
            else:
                f = np.full_like(f, 0.5)
            val = lo + (hi - lo) * f
            if c in (0, 2):
                out[:, m, c] = 1 - val
            else:
                out[:, m, c] = val
    return out


def thinfilm_morphology(n=2048, seed=1, period=40.0, device="cpu", dtype=torch.float32):
    """3-phase 2D spinodal-like film (n x n x 1): Gaussian random field with
    spectral peak at |q| = 2 pi/period (in voxels), thresholded at terciles with a
    2-voxel smooth interface; in-plane director tangent to the phase boundaries,
    S = 0.8 at interfaces and 0.2 in the bulk."""
    noise = _randn((n, n), seed, (0, 0), device)
    spec = torch.fft.fft2(noise)
    f = torch.fft.fftfreq(n, device=device)
    fy, fx = torch.meshgrid(f, f, indexing="ij")
    q = torch.sqrt(fx * fx + fy * fy)
    q0 = 1.0 / period
    filt = torch.exp(-((q - q0) ** 2) / (2 * (0.25 * q0) ** 2))
    phi = torch.fft.ifft2(spec * filt).real
    phi = _standardise(phi)
    flat = phi.reshape(-1)
    t1 = torch.quantile(flat[:: max(1, flat.numel() // 1_000_000)], 1 / 3)
    t2 = torch.quantile(flat[:: max(1, flat.numel() // 1_000_000)], 2 / 3)
    width = 2.0 * float(phi.std()) / period * 2.0  # about a 2-voxel interface
    s1 = torch.sigmoid((phi - t1) / width)
    s2 = torch.sigmoid((phi - t2) / width)
    v0 = 1 - s1
    v1 = s1 - s2
    v2 = s2
    gy = (torch.roll(phi, -1, 0) - torch.roll(phi, 1, 0)) * 0.5
    gx = (torch.roll(phi, -1, 1) - torch.roll(phi, 1, 1)) * 0.5
    psi = torch.atan2(gy, gx) + math.pi / 2
    iface = (s1 * (1 - s1) + s2 * (1 - s2)) * 4.0  # ~1 at interfaces, ~0 in bulk
    S = 0.2 + 0.6 * torch.clamp(iface, 0, 1)
    V = n * n
    out = torch.empty((3, V, 4), dtype=dtype, device=device)
    for m, vf in enumerate((v0, v1, v2)):
        out[m, :, 0] = S.reshape(V).to(dtype)
        out[m, :, 1] = math.pi / 2
        out[m, :, 2] = psi.reshape(V).to(dtype)
        out[m, :, 3] = vf.reshape(V).to(dtype)
    return out


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def config0():
    """64^3 sphere, 2 materials, PhysSize 5 nm, 284/285/286 eV, EAngle 0:30:150, NN."""
    energies = [284.0, 285.0, 286.0]
    return dict(dims=(64, 64, 64), num_materials=2, phys_size=5.0, energies=energies,
                angles=(0.0, 30.0, 150.0), interpolation=0, windowing=0,
                voxels=sphere_morphology(64, 20.0), mats=sphere_constants(energies))


def config1_desc():
    energies = [280.0 + i for i in range(20)]
    return dict(dims=(512, 512, 128), num_materials=4, phys_size=2.0, energies=energies,
                angles=(0.0, 5.0, 175.0), interpolation=1, windowing=1)


def config2_desc():
    energies = [280.0 + 0.5 * i for i in range(50)]
    return dict(dims=(2048, 2048, 1), num_materials=3, phys_size=1.0, energies=energies,
                angles=(0.0, 2.0, 358.0), interpolation=1, windowing=0)


BATCH_GRIDS = ((128, 128, 32), (256, 256, 64), (384, 384, 48), (512, 512, 64))


def config3_members(count=64, seed=1):
    """Per-member (dims, M, energies) for the uneven batched launch."""
    rng = np.random.default_rng(seed)
    members = []
    for i in range(count):
        dims = BATCH_GRIDS[int(rng.integers(0, len(BATCH_GRIDS)))]
        M = int(rng.integers(2, 5))
        nE = int(rng.integers(8, 33))
        energies = list(np.linspace(280.0, 300.0, nE))
        members.append(dict(dims=dims, num_materials=M, phys_size=2.0, energies=energies,
                            angles=(0.0, 10.0, 170.0), interpolation=1, windowing=0,
                            seed=seed * 1000 + i))
    return members
