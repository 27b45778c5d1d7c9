"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (cyrsoxs_amd) never
imports this module.

Array conventions follow cyrsoxs_oracle.h: morphology float64[M][V][4],
optical constants float64[nE][M][4] = (npara.re, npara.im, nperp.re, nperp.im),
complex volumes complex128[V] (X fastest), patterns float64[nE][nK][Ny][Nx].
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class OrcDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
        ("num_materials", ctypes.c_uint32),
        ("phys_size", ctypes.c_double),
        ("num_energies", ctypes.c_uint32),
        ("energies", _dp),
        ("angle_start", ctypes.c_double), ("angle_increment", ctypes.c_double),
        ("angle_end", ctypes.c_double),
        ("num_kvectors", ctypes.c_uint32),
        ("kvectors", _dp),
        ("detector", ctypes.c_double * 3),
        ("case_type", ctypes.c_int32),
        ("morphology_type", ctypes.c_int32),
        ("reference_frame", ctypes.c_int32),
        ("interpolation", ctypes.c_int32),
        ("windowing", ctypes.c_int32),
        ("rot_mask", ctypes.c_int32),
        ("scatter_approach", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        u32 = ctypes.c_uint32
        d = ctypes.c_double
        sig = {
            "orc_rotation_vec_to_vec": [vp, vp, vp],
            "orc_rodrigues": [vp, vp, vp, d],
            "orc_rotation_matrix_k": [vp, vp],
            "orc_base_configuration": [vp, vp, vp, vp],
            "orc_rotation_matrix_e": [vp, vp, d, vp],
            "orc_warp_affine_matrix": [vp, vp, vp],
            "orc_polarization": [vp, vp, vp, vp, vp, vp, vp],
            "orc_compute_nt": [vp, vp, vp, vp],
            "orc_polarization_nt": [vp, vp, vp, vp, vp, vp],
            "orc_fft3d": [vp, u32, u32, u32],
            "orc_fft1d": [vp, u32],
            "orc_dft1d_direct": [vp, vp, u32],
            "orc_dc_replace": [vp, u32, u32, u32],
            "orc_igor_shift": [vp, u32, u32, u32],
            "orc_scatter3d": [vp, vp, vp, vp, d, vp, vp],
            "orc_ewald_partial": [vp, vp, vp, vp, d, vp, vp],
            "orc_ewald_full": [vp, vp, d, vp, vp],
            "orc_warp_affine": [vp, vp, u32, u32, vp],
            "orc_rotation_coeffs": [d, u32, u32, vp],
            "orc_simulate": [vp, vp, vp, vp],
            "orc_simulate_subset": [vp, vp, vp, vp, u32, u32, u32, u32],
            "orc_set_threads": [ctypes.c_int],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None
        L.orc_num_angles.argtypes = [d, d, d]
        L.orc_num_angles.restype = u32
        L.orc_simulate.restype = ctypes.c_int
        L.orc_simulate_subset.restype = ctypes.c_int
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Sim:
    """Simulation description mirrored into the oracle's orc_desc."""

    def __init__(self, dims, num_materials, phys_size, energies, angles, kvectors=((0, 0, 1),),
                 detector=(0, 0, 1), case_type=0, morphology_type=0, reference_frame=1,
                 interpolation=1, windowing=0, rot_mask=0, scatter_approach=0, algorithm=0):
        self.dims = tuple(int(x) for x in dims)  # (Nx, Ny, Nz)
        self.num_materials = int(num_materials)
        self.phys_size = float(phys_size)
        self.energies = _f64(energies)
        self.angles = tuple(float(x) for x in angles)  # (start, increment, end) degrees
        self.kvectors = _f64(kvectors, (-1, 3))
        self.detector = tuple(float(x) for x in detector)
        self.case_type = case_type
        self.morphology_type = morphology_type
        self.reference_frame = reference_frame
        self.interpolation = interpolation
        self.windowing = windowing
        self.rot_mask = rot_mask
        self.scatter_approach = scatter_approach
        self.algorithm = algorithm

    def desc(self):
        d = OrcDesc()
        d.nx, d.ny, d.nz = self.dims
        d.num_materials = self.num_materials
        d.phys_size = self.phys_size
        d.num_energies = len(self.energies)
        d.energies = self.energies.ctypes.data_as(_dp)
        d.angle_start, d.angle_increment, d.angle_end = self.angles
        d.num_kvectors = len(self.kvectors)
        d.kvectors = self.kvectors.ctypes.data_as(_dp)
        for i in range(3):
            d.detector[i] = self.detector[i]
        d.case_type = self.case_type
        d.morphology_type = self.morphology_type
        d.reference_frame = self.reference_frame
        d.interpolation = self.interpolation
        d.windowing = self.windowing
        d.rot_mask = self.rot_mask
        d.scatter_approach = self.scatter_approach
        d.algorithm = self.algorithm
        return d

    @property
    def num_angles(self):
        return int(lib().orc_num_angles(*self.angles))

    @property
    def nvox(self):
        nx, ny, nz = self.dims
        return nx * ny * nz


def simulate(sim: Sim, voxels, mats, energies=None, angles=None):
    """Whole pipeline in FP64.  Returns float64[nE][nK][Ny][Nx]."""
    L = lib()
    vox = _f64(voxels, (sim.num_materials, sim.nvox, 4))
    m = _f64(mats, (len(sim.energies), sim.num_materials, 4))
    nx, ny, nz = sim.dims
    e0, e1 = (0, len(sim.energies)) if energies is None else energies
    a0, a1 = (0, sim.num_angles) if angles is None else angles
    out = np.empty((e1 - e0, len(sim.kvectors), ny, nx), dtype=np.float64)
    d = sim.desc()
    rc = L.orc_simulate_subset(ctypes.byref(d), _p(vox), _p(m), _p(out), e0, e1, a0, a1)
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out


def polarization(sim: Sim, voxels, mats_e, R):
    L = lib()
    vox = _f64(voxels, (sim.num_materials, sim.nvox, 4))
    m = _f64(mats_e, (sim.num_materials, 4))
    R = _f64(R, (9,))
    px = np.empty(sim.nvox, np.complex128)
    py = np.empty_like(px)
    pz = np.empty_like(px)
    d = sim.desc()
    L.orc_polarization(ctypes.byref(d), _p(vox), _p(m), _p(R), _p(px), _p(py), _p(pz))
    return px, py, pz


def compute_nt(sim: Sim, voxels, mats_e):
    L = lib()
    vox = _f64(voxels, (sim.num_materials, sim.nvox, 4))
    m = _f64(mats_e, (sim.num_materials, 4))
    nt = np.empty((6, sim.nvox), np.complex128)
    d = sim.desc()
    L.orc_compute_nt(ctypes.byref(d), _p(vox), _p(m), _p(nt))
    return nt


def polarization_nt(sim: Sim, nt, R):
    L = lib()
    nt = np.ascontiguousarray(nt, dtype=np.complex128)
    R = _f64(R, (9,))
    px = np.empty(sim.nvox, np.complex128)
    py = np.empty_like(px)
    pz = np.empty_like(px)
    d = sim.desc()
    L.orc_polarization_nt(ctypes.byref(d), _p(nt), _p(R), _p(px), _p(py), _p(pz))
    return px, py, pz


def fft3d(vol):
    """vol: complex array shaped (Nz, Ny, Nx); returns the forward unnormalised DFT."""
    a = np.array(vol, dtype=np.complex128, order="C", copy=True)
    nz, ny, nx = a.shape
    lib().orc_fft3d(_p(a), nx, ny, nz)
    return a


def fft1d(x):
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_fft1d(_p(a), len(a))
    return a


def dft1d_direct(x):
    a = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(a)
    lib().orc_dft1d_direct(_p(a), _p(out), len(a))
    return out


def dc_replace(vol):
    a = np.array(vol, dtype=np.complex128, order="C", copy=True)
    nz, ny, nx = a.shape
    lib().orc_dc_replace(_p(a), nx, ny, nz)
    return a


def igor_shift(vol):
    a = np.array(vol, dtype=np.complex128, order="C", copy=True)
    nz, ny, nx = a.shape
    lib().orc_igor_shift(_p(a), nx, ny, nz)
    return a


def ewald_partial(sim: Sim, px, py, pz, k, kvec=(0, 0, 1)):
    nx, ny, nz = sim.dims
    kv = _f64(kvec, (3,))
    arrs = [np.ascontiguousarray(p, dtype=np.complex128).reshape(-1) for p in (px, py, pz)]
    out = np.empty(nx * ny, np.float64)
    d = sim.desc()
    lib().orc_ewald_partial(ctypes.byref(d), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), k, _p(kv), _p(out))
    return out.reshape(ny, nx)


def scatter3d(sim: Sim, px, py, pz, k, kvec=(0, 0, 1)):
    kv = _f64(kvec, (3,))
    arrs = [np.ascontiguousarray(p, dtype=np.complex128).reshape(-1) for p in (px, py, pz)]
    out = np.empty(sim.nvox, np.float64)
    d = sim.desc()
    lib().orc_scatter3d(ctypes.byref(d), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), k, _p(kv), _p(out))
    return out


def ewald_full(sim: Sim, scatter, k, kvec=(0, 0, 1)):
    nx, ny, nz = sim.dims
    kv = _f64(kvec, (3,))
    s = _f64(scatter).reshape(-1)
    out = np.empty(nx * ny, np.float64)
    d = sim.desc()
    lib().orc_ewald_full(ctypes.byref(d), _p(s), k, _p(kv), _p(out))
    return out.reshape(ny, nx)


def warp_affine(src, coeffs, prefill=np.nan):
    s = _f64(src)
    ny, nx = s.shape
    dst = np.full_like(s, prefill)
    c = _f64(coeffs, (6,))
    lib().orc_warp_affine(_p(s), _p(dst), nx, ny, _p(c))
    return dst


def rotation_coeffs(angle_rad, nx, ny):
    c = np.empty(6, np.float64)
    lib().orc_rotation_coeffs(angle_rad, nx, ny, _p(c))
    return c


def rotation_vec_to_vec(a, b):
    R = np.empty(9, np.float64)
    lib().orc_rotation_vec_to_vec(_p(_f64(a, (3,))), _p(_f64(b, (3,))), _p(R))
    return R.reshape(3, 3)


def rodrigues(v, axis, angle):
    out = np.empty(3, np.float64)
    lib().orc_rodrigues(_p(out), _p(_f64(v, (3,))), _p(_f64(axis, (3,))), angle)
    return out


def rotation_matrix_k(k):
    R = np.empty(9, np.float64)
    lib().orc_rotation_matrix_k(_p(_f64(k, (3,))), _p(R))
    return R.reshape(3, 3)


def base_configuration(k):
    RK = rotation_matrix_k(k).reshape(-1).copy()
    R = np.empty(9, np.float64)
    ang = np.zeros(1, np.float64)
    lib().orc_base_configuration(_p(_f64(k, (3,))), _p(RK), _p(R), _p(ang))
    return R.reshape(3, 3), float(ang[0])


def rotation_matrix_e(k, angle):
    RK = rotation_matrix_k(k).reshape(-1).copy()
    R = np.empty(9, np.float64)
    lib().orc_rotation_matrix_e(_p(_f64(k, (3,))), _p(RK), angle, _p(R))
    return R.reshape(3, 3)


def warp_affine_matrix(src_pts, dst_pts):
    c = np.empty(6, np.float64)
    lib().orc_warp_affine_matrix(_p(_f64(src_pts, (6,))), _p(_f64(dst_pts, (6,))), _p(c))
    return c


def num_angles(start, inc, end):
    return int(lib().orc_num_angles(start, inc, end))


def max_threads():
    return int(lib().orc_max_threads())


def set_threads(n):
    lib().orc_set_threads(int(n))
