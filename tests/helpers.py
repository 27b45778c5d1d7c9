"""Shared test utilities: comparison metrics and small seeded morphologies."""
import numpy as np


def rel_l2(a, b, mask=None):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    if mask is not None:
        a = a[mask]
        b = b[mask]
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def pattern_rel_l2(gpu, ref):
    """Per-pattern relative L2 over the detector mask (pixels finite in ref).
    Returns (max over patterns, nan-mask-identical flag)."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert gpu.shape == ref.shape, (gpu.shape, ref.shape)
    same_mask = bool(np.array_equal(np.isnan(gpu), np.isnan(ref)))
    worst = 0.0
    flat_g = gpu.reshape(-1, gpu.shape[-2], gpu.shape[-1])
    flat_r = ref.reshape(-1, ref.shape[-2], ref.shape[-1])
    for g, r in zip(flat_g, flat_r):
        m = np.isfinite(r)
        if m.any():
            worst = max(worst, rel_l2(g, r, m))
    return worst, same_mask


def random_euler(dims, M, seed, s_zero_frac=0.1):
    """[M][V][4] Euler morphology: Vfrac from a Dirichlet draw, S uniform (some 0),
    theta/psi uniform; theta = psi = 0 where S == 0 (reader rule)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    V = nx * ny * nz
    vf = rng.dirichlet(np.ones(M), size=V).T
    out = np.zeros((M, V, 4))
    for m in range(M):
        S = rng.uniform(0, 1, V)
        S[rng.uniform(0, 1, V) < s_zero_frac] = 0
        th = np.arccos(1 - 2 * rng.uniform(0, 1, V))
        ps = 2 * np.pi * rng.uniform(0, 1, V)
        th[S == 0] = 0
        ps[S == 0] = 0
        out[m, :, 0] = S
        out[m, :, 1] = th
        out[m, :, 2] = ps
        out[m, :, 3] = vf[m]
    return out


def random_vector(dims, M, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    V = nx * ny * nz
    out = np.zeros((M, V, 4))
    for m in range(M):
        s = rng.normal(size=(V, 3)) * 0.3
        out[m, :, :3] = s
        out[m, :, 3] = rng.uniform(0, 0.5, V)
    return out


def random_constants(nE, M, seed):
    rng = np.random.default_rng(seed)
    out = np.zeros((nE, M, 4))
    out[:, :, 0] = 1 - rng.uniform(1e-4, 3e-3, (nE, M))
    out[:, :, 1] = rng.uniform(1e-4, 3e-3, (nE, M))
    out[:, :, 2] = 1 - rng.uniform(1e-4, 3e-3, (nE, M))
    out[:, :, 3] = rng.uniform(1e-4, 3e-3, (nE, M))
    return out
