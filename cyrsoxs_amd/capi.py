"""ctypes binding of the C ABI in include/cyrsoxs.h (lib/libcyrsoxs.so).

This is the Python view of the drop-in boundary: plain pointers and POD
descriptors.  Device buffers are torch tensors (PyTorch supplies HBM
allocations and streams; all compute is in the HIP library).  Loading fails
loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcyrsoxs.so")

CYR_OK = 0
STATUS = {0: "CYR_OK", 1: "CYR_E_INVALID", 2: "CYR_E_NOMEM", 3: "CYR_E_DEVICE", 4: "CYR_E_ADMISSION",
          5: "CYR_E_CANCELLED", 6: "CYR_E_UNSUPPORTED", 7: "CYR_E_NODEVICE", 8: "CYR_E_TIMEOUT",
          9: "CYR_E_NOTFOUND"}
FP32, FP64 = 0, 1
STORE_NATIVE, STORE_FP16, STORE_BF16 = 0, 1, 2
PIPE_PLAIN, PIPE_EBASIS, PIPE_MBASIS, PIPE_AUTO = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)


class SimDesc(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_uint32 * 3),
        ("num_materials", ctypes.c_uint32),
        ("phys_size", ctypes.c_double),
        ("num_energies", ctypes.c_uint32),
        ("energies", _dp),
        ("angle_start", ctypes.c_double),
        ("angle_increment", ctypes.c_double),
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
        ("precision", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("pipeline", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class EngineStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("resident_bytes", ctypes.c_uint64),
                ("budget_bytes", ctypes.c_uint64), ("in_flight", ctypes.c_uint32), ("queued", ctypes.c_uint32)]


class CyrError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)} {detail}".strip())


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "cyr_simulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_simulate_device": (ctypes.c_int, [ctypes.c_void_p] * 5),
    "cyr_simulate_device_noalloc": (ctypes.c_int, [ctypes.c_void_p] * 5),
    "cyr_polarization": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                        ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_polarization_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_fft3d_device": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "cyr_fft_axis_device": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                           ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p]),
    "cyr_dc_shift_device": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                           ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]),
    "cyr_ewald_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_warp_affine_device": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_simulate_batch": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_engine_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p]),
    "cyr_engine_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_engine_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p]),
    "cyr_engine_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_engine_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "cyr_engine_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_engine_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cyr_engine_request_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "cyr_engine_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "cyr_batch_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "cyr_batch_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cyr_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "cyr_last_error": (ctypes.c_char_p, []),
    "cyr_device_count": (ctypes.c_int, []),
    "cyr_num_angles": (ctypes.c_uint32, [ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "cyr_footprint_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
    "cyr_resolved_pipeline": (ctypes.c_int32, [ctypes.c_void_p]),
    "cyr_release_workspace": (ctypes.c_int, []),
    "cyr_version": (ctypes.c_char_p, []),
    "cyr_kernel_timing": (ctypes.c_int, [ctypes.c_int]),
    "cyr_set_tuning": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32, ctypes.c_void_p]),
    "cyr_get_tuning": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32, ctypes.c_void_p]),
    "cyr_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "cyr_set_energy_lanes": (ctypes.c_int, [ctypes.c_uint32]),
    "cyr_energy_lanes": (ctypes.c_uint32, []),
    "cyr_kernel_timing_report": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
}


def exported_symbols():
    return sorted(_SIGS)


def lib():
    """Loads lib/libcyrsoxs.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m cyrsoxs_amd.build` "
                              "(the HIP library is the only implementation)")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != CYR_OK:
        detail = lib().cyr_last_error()
        raise CyrError(status, where, detail.decode() if detail else "")


class Sim:
    """Keeps a SimDesc and the arrays it points to alive together."""

    def __init__(self, dims, num_materials, phys_size, energies, angles, kvectors=None, detector=(0, 0, 1),
                 case_type=0, morphology_type=0, reference_frame=1, interpolation=1, windowing=0, rot_mask=0,
                 scatter_approach=0, algorithm=0, precision=FP32, storage=STORE_NATIVE, pipeline=PIPE_PLAIN):
        self.energies = np.ascontiguousarray(energies, dtype=np.float64)
        kv = np.zeros((0, 3)) if kvectors is None else np.asarray(kvectors, dtype=np.float64).reshape(-1, 3)
        self.kvectors = np.ascontiguousarray(kv)
        d = SimDesc()
        d.dims[0], d.dims[1], d.dims[2] = (int(x) for x in dims)
        d.num_materials = int(num_materials)
        d.phys_size = float(phys_size)
        d.num_energies = len(self.energies)
        d.energies = self.energies.ctypes.data_as(_dp)
        d.angle_start, d.angle_increment, d.angle_end = (float(x) for x in angles)
        d.num_kvectors = len(self.kvectors)
        d.kvectors = self.kvectors.ctypes.data_as(_dp) if len(self.kvectors) else None
        for i in range(3):
            d.detector[i] = float(detector[i])
        d.case_type = case_type
        d.morphology_type = morphology_type
        d.reference_frame = reference_frame
        d.interpolation = interpolation
        d.windowing = windowing
        d.rot_mask = rot_mask
        d.scatter_approach = scatter_approach
        d.algorithm = algorithm
        d.precision = precision
        d.storage = storage
        d.pipeline = pipeline
        self.desc = d

    @property
    def ref(self):
        return ctypes.byref(self.desc)

    @property
    def dims(self):
        return tuple(self.desc.dims)

    @property
    def num_angles(self):
        return int(lib().cyr_num_angles(self.desc.angle_start, self.desc.angle_increment, self.desc.angle_end))

    @property
    def num_k(self):
        return max(1, len(self.kvectors))

    @property
    def real_dtype(self):
        return np.float64 if self.desc.precision == FP64 else np.float32

    def footprint(self):
        return int(lib().cyr_footprint_bytes(self.ref))

    def resolved_pipeline(self):
        """The executor this simulation runs (PIPE_AUTO resolved)."""
        return int(lib().cyr_resolved_pipeline(self.ref))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# The C side receives bare pointers and cannot check them: every buffer is
# validated here (element count, element type, contiguity, device) before a
# pointer crosses the boundary.

def _bad(where, msg):
    raise CyrError(1, where, msg)  # CYR_E_INVALID


def _voxel_count(sim):
    nx, ny, nz = sim.dims
    return int(sim.desc.num_materials) * nx * ny * nz * 4


def _pattern_count(sim):
    nx, ny, _ = sim.dims
    return len(sim.energies) * sim.num_k * ny * nx


def _mats(sim, mats, where):
    m = np.ascontiguousarray(mats, dtype=np.float64)
    want = len(sim.energies) * int(sim.desc.num_materials) * 4
    if m.size != want:
        _bad(where, f"optical constants hold {m.size} values; [nE][M][4] = {want} expected")
    return m


def _host_voxels(sim, voxels, where):
    vox = np.ascontiguousarray(voxels, dtype=sim.real_dtype)
    if vox.size != _voxel_count(sim):
        _bad(where, f"morphology holds {vox.size} values; [M][V][4] = {_voxel_count(sim)} expected")
    return vox


def _host_out(sim, out, where):
    if out is None:
        nx, ny, _ = sim.dims
        return np.empty((len(sim.energies), sim.num_k, ny, nx), dtype=sim.real_dtype)
    if not isinstance(out, np.ndarray) or out.dtype != sim.real_dtype or not out.flags.c_contiguous \
            or not out.flags.writeable or out.size != _pattern_count(sim):
        _bad(where, f"output must be a writeable C-contiguous {np.dtype(sim.real_dtype).name} array of "
                    f"{_pattern_count(sim)} values")
    return out


def _storage_dtypes(sim):
    import torch
    st = int(sim.desc.storage)
    if st == STORE_FP16:
        return (torch.float16,)
    if st == STORE_BF16:
        return (torch.bfloat16,)
    return (torch.float64,) if sim.desc.precision == FP64 else (torch.float32,)


def _device_tensor(t, dtypes, count, what, where):
    import torch
    if not isinstance(t, torch.Tensor):
        _bad(where, f"{what} must be a torch tensor on the GPU")
    if t.dtype not in dtypes:
        _bad(where, f"{what} has dtype {t.dtype}; {' or '.join(str(d) for d in dtypes)} expected")
    if not t.is_contiguous():
        _bad(where, f"{what} is not contiguous")
    if t.numel() != count:
        _bad(where, f"{what} holds {t.numel()} elements; {count} expected")


def _device_args(sim, d_voxels, d_out, where):
    import torch
    _device_tensor(d_voxels, _storage_dtypes(sim), _voxel_count(sim), "morphology", where)
    od = torch.float64 if sim.desc.precision == FP64 else torch.float32
    _device_tensor(d_out, (od,), _pattern_count(sim), "output", where)
    for t, what in ((d_voxels, "morphology"), (d_out, "output")):
        if not t.is_cuda:
            _bad(where, f"{what} is on {t.device}, not on a GPU")
    if d_voxels.device != d_out.device:
        _bad(where, f"morphology on {d_voxels.device} and output on {d_out.device}")


def simulate(sim: Sim, voxels, mats):
    """Host API: voxels [M][V][4] (compute precision), mats [nE][M][4] -> patterns [nE][nK][Ny][Nx]."""
    vox = _host_voxels(sim, voxels, "cyr_simulate")
    m = _mats(sim, mats, "cyr_simulate")
    out = _host_out(sim, None, "cyr_simulate")
    check(lib().cyr_simulate(sim.ref, _ptr(vox), _ptr(m), _ptr(out)), "cyr_simulate")
    return out


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def simulate_device(sim: Sim, d_voxels, mats, d_out, stream=None):
    """Device API over torch tensors (inputs resident in HBM)."""
    m = _mats(sim, mats, "cyr_simulate_device")
    _device_args(sim, d_voxels, d_out, "cyr_simulate_device")
    check(lib().cyr_simulate_device(sim.ref, ctypes.c_void_p(d_voxels.data_ptr()), _ptr(m),
                                    ctypes.c_void_p(d_out.data_ptr()), _stream_ptr(stream)), "cyr_simulate_device")


def simulate_device_noalloc(sim: Sim, d_voxels, mats, d_out, stream=None):
    """cyr_simulate_device_noalloc: the device API with every work buffer in the
    library's static scratch array (no device allocation); simulations larger
    than NOALLOC_SCRATCH_BYTES are rejected with CYR_E_INVALID."""
    m = _mats(sim, mats, "cyr_simulate_device_noalloc")
    _device_args(sim, d_voxels, d_out, "cyr_simulate_device_noalloc")
    check(lib().cyr_simulate_device_noalloc(sim.ref, ctypes.c_void_p(d_voxels.data_ptr()), _ptr(m),
                                            ctypes.c_void_p(d_out.data_ptr()), _stream_ptr(stream)),
          "cyr_simulate_device_noalloc")


NOALLOC_SCRATCH_BYTES = 128 * 1024 * 1024  # CYR_NOALLOC_SCRATCH_BYTES


class YzTuning(ctypes.Structure):
    _fields_ = [("columns", ctypes.c_uint32), ("target_blocks", ctypes.c_uint32)]


def autotune(sim: Sim):
    """Times the E-basis Y/z launch shapes for sim's grid and records the fastest;
    returns (columns, target_blocks) ((0, 0): nothing to tune)."""
    t = YzTuning()
    check(lib().cyr_autotune(sim.ref, ctypes.byref(t)), "cyr_autotune")
    return int(t.columns), int(t.target_blocks)


def set_energy_lanes(lanes):
    check(lib().cyr_set_energy_lanes(int(lanes)), "cyr_set_energy_lanes")


def energy_lanes():
    return int(lib().cyr_energy_lanes())


def set_tuning(dims, precision, columns, target_blocks):
    d = (ctypes.c_uint32 * 3)(*[int(x) for x in dims])
    t = YzTuning(int(columns), int(target_blocks))
    check(lib().cyr_set_tuning(d, int(precision), ctypes.byref(t)), "cyr_set_tuning")


def get_tuning(dims, precision):
    d = (ctypes.c_uint32 * 3)(*[int(x) for x in dims])
    t = YzTuning()
    check(lib().cyr_get_tuning(d, int(precision), ctypes.byref(t)), "cyr_get_tuning")
    return int(t.columns), int(t.target_blocks)


def polarization(sim: Sim, voxels, mats, energy_id, eangle_rad):
    rd = sim.real_dtype
    cd = np.complex128 if rd == np.float64 else np.complex64
    vox = _host_voxels(sim, voxels, "cyr_polarization")
    m = _mats(sim, mats, "cyr_polarization")
    nx, ny, nz = sim.dims
    V = nx * ny * nz
    px, py, pz = (np.empty(V, dtype=cd) for _ in range(3))
    check(lib().cyr_polarization(sim.ref, _ptr(vox), _ptr(m), int(energy_id), float(eangle_rad), _ptr(px), _ptr(py),
                                 _ptr(pz)), "cyr_polarization")
    return px, py, pz


def fft3d_device(vols, dims, precision, stream=None):
    arr = (ctypes.c_void_p * len(vols))(*[v.data_ptr() for v in vols])
    nx, ny, nz = dims
    check(lib().cyr_fft3d_device(nx, ny, nz, precision, arr, len(vols), _stream_ptr(stream)), "cyr_fft3d_device")


def fft_axis_device(vols, dims, precision, axis, stream=None):
    """One axis pass, queued asynchronously on the stream (no synchronisation)."""
    arr = (ctypes.c_void_p * len(vols))(*[v.data_ptr() for v in vols])
    nx, ny, nz = dims
    check(lib().cyr_fft_axis_device(nx, ny, nz, precision, arr, len(vols), axis, _stream_ptr(stream)),
          "cyr_fft_axis_device")


def dc_shift_device(vol, dims, precision, do_dc=True, do_shift=True, stream=None):
    nx, ny, nz = dims
    check(lib().cyr_dc_shift_device(nx, ny, nz, precision, ctypes.c_void_p(vol.data_ptr()), int(do_dc),
                                    int(do_shift), _stream_ptr(stream)), "cyr_dc_shift_device")


def ewald_device(sim: Sim, energy_id, k_id, px, py, pz, proj, stream=None):
    check(lib().cyr_ewald_device(sim.ref, energy_id, k_id, ctypes.c_void_p(px.data_ptr()),
                                 ctypes.c_void_p(py.data_ptr()), ctypes.c_void_p(pz.data_ptr()),
                                 ctypes.c_void_p(proj.data_ptr()), _stream_ptr(stream)), "cyr_ewald_device")


def warp_affine_device(src, dst, coeffs, precision, stream=None):
    ny, nx = src.shape[-2:]
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    check(lib().cyr_warp_affine_device(nx, ny, precision, ctypes.c_void_p(src.data_ptr()),
                                       ctypes.c_void_p(dst.data_ptr()), _ptr(c), _stream_ptr(stream)),
          "cyr_warp_affine_device")


def _batch_args(sims, d_voxels, mats, d_outs, where):
    if not (len(sims) == len(d_voxels) == len(mats) == len(d_outs)) or not sims:
        _bad(where, "sims, morphologies, constants and outputs must be non-empty lists of one length")
    for s, v, o in zip(sims, d_voxels, d_outs):
        _device_args(s, v, o, where)
    return [_mats(s, m, where) for s, m in zip(sims, mats)]


def simulate_batch(sims, d_voxels, mats, d_outs, stream=None):
    n = len(sims)
    mats_c = _batch_args(sims, d_voxels, mats, d_outs, "cyr_simulate_batch")
    descs = (SimDesc * n)(*[s.desc for s in sims])
    vp = (ctypes.c_void_p * n)(*[v.data_ptr() for v in d_voxels])
    mp = (ctypes.c_void_p * n)(*[m.ctypes.data for m in mats_c])
    op = (ctypes.c_void_p * n)(*[o.data_ptr() for o in d_outs])
    check(lib().cyr_simulate_batch(n, descs, vp, mp, op, _stream_ptr(stream)), "cyr_simulate_batch")


class BatchPlan:
    """A batch captured once in a hipGraph (cyr_batch_create) and replayed with
    launch(); the torch tensors given here are bound for its lifetime."""

    def __init__(self, sims, d_voxels, mats, d_outs, lanes=0):
        n = len(sims)
        self.handle = ctypes.c_void_p()
        mats_c = _batch_args(sims, d_voxels, mats, d_outs, "cyr_batch_create")
        self._keep = (sims, d_voxels, d_outs)
        descs = (SimDesc * n)(*[s.desc for s in sims])
        vp = (ctypes.c_void_p * n)(*[v.data_ptr() for v in d_voxels])
        mp = (ctypes.c_void_p * n)(*[m.ctypes.data for m in mats_c])
        op = (ctypes.c_void_p * n)(*[o.data_ptr() for o in d_outs])
        check(lib().cyr_batch_create(n, descs, vp, mp, op, int(lanes), ctypes.byref(self.handle)), "cyr_batch_create")

    def launch(self, stream=None):
        check(lib().cyr_batch_launch(self.handle, _stream_ptr(stream)), "cyr_batch_launch")

    def close(self):
        if self.handle:
            check(lib().cyr_batch_destroy(self.handle), "cyr_batch_destroy")
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EngineStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("resident_bytes", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("in_flight", ctypes.c_uint32), ("queued", ctypes.c_uint32)]


class Engine:
    """Serving engine (cyr_engine_*): submit() returns a request id; host inputs
    and outputs are numpy arrays kept alive here until the request is released."""

    def __init__(self, device=0, budget_bytes=0, slots=4):
        self.handle = ctypes.c_void_p()
        check(lib().cyr_engine_create(int(device), int(budget_bytes), int(slots), ctypes.byref(self.handle)),
              "cyr_engine_create")
        self._live = {}

    def submit(self, sim: Sim, voxels, mats, out=None):
        vox = _host_voxels(sim, voxels, "cyr_engine_submit")
        m = _mats(sim, mats, "cyr_engine_submit")
        out = _host_out(sim, out, "cyr_engine_submit")
        rid = ctypes.c_uint64()
        check(lib().cyr_engine_submit(self.handle, sim.ref, _ptr(vox), _ptr(m), _ptr(out), ctypes.byref(rid)),
              "cyr_engine_submit")
        self._live[rid.value] = (sim, vox, m, out)
        return rid.value

    def wait(self, rid, timeout=-1.0):
        st = ctypes.c_int()
        check(lib().cyr_engine_wait(self.handle, ctypes.c_uint64(rid), float(timeout), ctypes.byref(st)),
              "cyr_engine_wait")
        return int(st.value)

    def poll(self, rid):
        done, st = ctypes.c_int(), ctypes.c_int()
        check(lib().cyr_engine_poll(self.handle, ctypes.c_uint64(rid), ctypes.byref(done), ctypes.byref(st)),
              "cyr_engine_poll")
        return bool(done.value), int(st.value)

    def result(self, rid):
        return self._live[rid][3]

    def times(self, rid):
        q, r, l = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().cyr_engine_request_times(self.handle, ctypes.c_uint64(rid), ctypes.byref(q), ctypes.byref(r),
                                             ctypes.byref(l)), "cyr_engine_request_times")
        return q.value, r.value, l.value

    def release(self, rid):
        check(lib().cyr_engine_release(self.handle, ctypes.c_uint64(rid)), "cyr_engine_release")
        return self._live.pop(rid)[3]

    def cancel(self, rid):
        check(lib().cyr_engine_cancel(self.handle, ctypes.c_uint64(rid)), "cyr_engine_cancel")

    def stats(self):
        s = EngineStats()
        check(lib().cyr_engine_get_stats(self.handle, ctypes.byref(s)), "cyr_engine_get_stats")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.handle:
            check(lib().cyr_engine_destroy(self.handle), "cyr_engine_destroy")
            self.handle = ctypes.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def kernel_timing(enable):
    check(lib().cyr_kernel_timing(1 if enable else 0), "cyr_kernel_timing")


def kernel_timing_report():
    """{kernel: {"calls", "total_us", "avg_us", "avg_bytes"}} accumulated since
    kernel_timing(True); avg_bytes is the launch's algorithmic HBM traffic."""
    buf = ctypes.create_string_buffer(1 << 16)
    lib().cyr_kernel_timing_report(buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        name, calls, tot, avg, byts = line.split()
        out[name] = {"calls": int(calls), "total_us": float(tot), "avg_us": float(avg), "avg_bytes": float(byts)}
    return out


def device_count():
    return int(lib().cyr_device_count())


def version():
    return lib().cyr_version().decode()
