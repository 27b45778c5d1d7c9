"""The C ABI library loads, exports every symbol include/cyrsoxs.h declares, and
its host-side validation behaves like the reference's input checks.  No compute
call needs a GPU here: every call below returns before touching a device (or
reports CYR_E_NODEVICE when there is none)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from cyrsoxs_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cyrsoxs.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cyr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (cyr_[a-z0-9_]+)", out))
    assert set(names) <= exported
    # the ctypes binding covers the whole ABI
    assert set(names) == set(capi.exported_symbols())


def test_status_strings_and_version():
    L = capi.lib()
    for code in range(10):
        assert L.cyr_status_string(code)
    assert b"gfx950" in L.cyr_version()


@pytest.mark.parametrize("tri,n", [((0, 30, 150), 6), ((0, 5, 175), 36), ((0, 2, 358), 180), ((0, 10, 170), 18),
                                   ((0, 0, 0), 1), ((10, 1, 10), 1)])
def test_num_angles(tri, n):
    assert capi.lib().cyr_num_angles(*map(float, tri)) == n


def _sim(**kw):
    base = dict(dims=(16, 16, 8), num_materials=2, phys_size=5.0, energies=[285.0], angles=(0, 30, 150))
    base.update(kw)
    return capi.Sim(**base)


BAD = {
    "start_gt_end": dict(angles=(10, 1, 0)),
    "zero_increment": dict(angles=(0, 0, 10)),
    "nx_lt_2": dict(dims=(1, 16, 8)),
    "no_materials": dict(num_materials=0),
    "negative_phys": dict(phys_size=-1.0),
    "no_energies": dict(energies=[]),
    "bad_energy": dict(energies=[float("nan")]),
    "bad_interp": dict(interpolation=2),
    "bad_window": dict(windowing=3),
    "bad_case": dict(case_type=3),
    "bad_morph": dict(morphology_type=2),
    "bad_precision": dict(precision=7),
    "bad_storage": dict(storage=5),
    "bad_pipeline": dict(pipeline=9),
    "case1_no_k": dict(case_type=1),
    "unnormalised_k": dict(case_type=1, kvectors=[[0, 0, 2.0]]),
}


@pytest.mark.parametrize("name", sorted(BAD))
def test_invalid_descriptors_rejected(name):
    s = _sim(**BAD[name])
    vox = np.zeros((2, 16 * 16 * 8, 4), np.float32)
    mats = np.zeros((max(1, len(s.energies)), 2, 4))
    out = np.zeros(16 * 16 * 10, np.float32)
    L = capi.lib()
    st = L.cyr_simulate(s.ref, ctypes.c_void_p(vox.ctypes.data), ctypes.c_void_p(mats.ctypes.data),
                        ctypes.c_void_p(out.ctypes.data))
    assert st == 1, (name, st)
    assert L.cyr_last_error()
    assert L.cyr_footprint_bytes(s.ref) == 0


def test_null_buffers_rejected():
    s = _sim()
    L = capi.lib()
    assert L.cyr_simulate(s.ref, None, None, None) == 1
    assert L.cyr_fft3d_device(8, 8, 8, 0, None, 1, None) == 1
    assert L.cyr_warp_affine_device(8, 8, 0, None, None, None, None) == 1


def test_start_equals_end_means_one_angle():
    s = _sim(angles=(30, 0, 30))
    assert s.num_angles == 1
    assert capi.lib().cyr_footprint_bytes(s.ref) > 0


def test_footprint_accounts_for_pipeline():
    plain = _sim(dims=(512, 512, 128), num_materials=4, pipeline=capi.PIPE_PLAIN)
    eb = _sim(dims=(512, 512, 128), num_materials=4, pipeline=capi.PIPE_EBASIS)
    V = 512 * 512 * 128
    fp, fe = plain.footprint(), eb.footprint()
    assert fp >= V * (16 * 4 + 24)  # morphology + 3 complex64 volumes
    assert fe >= V * (16 * 4 + 40)  # morphology + 5 complex64 tensor fields
    assert fe > fp
    half = _sim(dims=(512, 512, 128), num_materials=4, storage=capi.STORE_FP16)
    assert plain.footprint() - half.footprint() == V * 4 * 8


def test_footprint_material_basis():
    """Material basis at 512^2 x 128, 4 materials: one pass of 7 packed complex64
    volumes plus band values, besides the morphology."""
    mb = _sim(dims=(512, 512, 128), num_materials=4, pipeline=capi.PIPE_MBASIS)
    V = 512 * 512 * 128
    assert mb.footprint() >= V * (16 * 4 + 7 * 8)


def test_auto_pipeline_resolution():
    """AUTO runs the E-basis executor for a single energy and the material basis for
    many; explicit choices are kept; grids without fused plans run plain."""
    one = _sim(dims=(512, 512, 128), num_materials=4, pipeline=capi.PIPE_AUTO, energies=[285.0])
    many = _sim(dims=(512, 512, 128), num_materials=4, pipeline=capi.PIPE_AUTO,
                energies=list(np.linspace(280, 299, 20)))
    assert one.resolved_pipeline() == capi.PIPE_EBASIS
    assert many.resolved_pipeline() == capi.PIPE_MBASIS
    for p in (capi.PIPE_PLAIN, capi.PIPE_EBASIS, capi.PIPE_MBASIS):
        assert _sim(dims=(64, 64, 32), num_materials=2, pipeline=p).resolved_pipeline() == p
    # 7 = a prime without a Stockham plan: every request runs the plain executor
    assert _sim(dims=(7, 7, 7), num_materials=2, pipeline=capi.PIPE_MBASIS).resolved_pipeline() == capi.PIPE_PLAIN
    bad = _sim(dims=(64, 64, 32), num_materials=2, pipeline=capi.PIPE_AUTO)
    bad.desc.pipeline = 9
    assert bad.resolved_pipeline() == -1


def test_no_device_reported_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    s = _sim()
    vox = np.zeros((2, 16 * 16 * 8, 4), np.float32)
    mats = np.zeros((1, 2, 4))
    out = np.zeros((1, 1, 16, 16), np.float32)
    with pytest.raises(capi.CyrError) as e:
        capi.simulate(s, vox, mats)
    assert e.value.status == 7


# ------------------------------------------------------------- argument checks
# The C side receives bare pointers; the binding rejects buffers of the wrong
# size, type, layout or device before any pointer crosses the boundary.

torch = pytest.importorskip("torch")


def _small():
    return _sim(dims=(8, 8, 4), num_materials=2, energies=[284.0, 285.0])


def _dev_bufs(sim, vdtype=torch.float32, odtype=torch.float32):
    nx, ny, nz = sim.dims
    vox = torch.zeros((2, nx * ny * nz, 4), dtype=vdtype)
    out = torch.zeros((2, 1, ny, nx), dtype=odtype)
    return vox, out


MATS = np.tile(np.array([1 - 1e-3, 1e-3, 1 - 1e-3, 1e-3]), (2, 2, 1))


@pytest.mark.parametrize("fn", ["simulate_device", "simulate_device_noalloc"])
def test_device_api_rejects_bad_buffers(fn):
    sim = _small()
    call = getattr(capi, fn)
    vox, out = _dev_bufs(sim)
    cases = {
        "numpy morphology": (vox.numpy(), out, "torch tensor"),
        "short morphology": (torch.zeros((2, 8 * 8 * 4 - 1, 4)), out, "elements"),
        "wrong morphology type": (vox.double(), out, "dtype"),
        "strided morphology": (torch.zeros((2, 4, 8 * 8 * 4)).transpose(1, 2), out, "contiguous"),
        "wrong output type": (vox, out.double(), "dtype"),
        "short output": (vox, out[:1], "elements"),
        "host tensors": (vox, out, "not on a GPU"),
    }
    for name, (v, o, msg) in cases.items():
        with pytest.raises(capi.CyrError) as e:
            call(sim, v, MATS, o)
        assert e.value.status == 1 and msg in str(e.value), (name, str(e.value))
    with pytest.raises(capi.CyrError) as e:
        call(sim, vox, MATS[:1], out)  # one energy's constants for two energies
    assert "optical constants" in str(e.value)


def test_storage_type_must_match_the_morphology_tensor():
    sim16 = _sim(dims=(8, 8, 4), num_materials=2, energies=[284.0, 285.0], storage=capi.STORE_FP16)
    vox, out = _dev_bufs(sim16)
    with pytest.raises(capi.CyrError) as e:
        capi.simulate_device(sim16, vox, MATS, out)  # float32 tensor for FP16 storage
    assert "float16" in str(e.value)
    vox16, _ = _dev_bufs(sim16, vdtype=torch.float16)
    with pytest.raises(capi.CyrError) as e:
        capi.simulate_device(sim16, vox16, MATS, out)  # right type, but on the host
    assert "not on a GPU" in str(e.value)
    sim64 = _sim(dims=(8, 8, 4), num_materials=2, energies=[284.0, 285.0], precision=capi.FP64)
    v64, o32 = _dev_bufs(sim64, vdtype=torch.float64)
    with pytest.raises(capi.CyrError) as e:
        capi.simulate_device(sim64, v64, MATS, o32)  # FP64 run writes float64 patterns
    assert "output" in str(e.value) and "dtype" in str(e.value)


def test_host_api_rejects_bad_sizes():
    sim = _small()
    with pytest.raises(capi.CyrError) as e:
        capi.simulate(sim, np.zeros((2, 8 * 8 * 4 - 1, 4)), MATS)
    assert "morphology holds" in str(e.value)
    with pytest.raises(capi.CyrError) as e:
        capi.simulate(sim, np.zeros((2, 8 * 8 * 4, 4)), MATS[:, :1])
    assert "optical constants" in str(e.value)
    with pytest.raises(capi.CyrError) as e:
        capi.polarization(sim, np.zeros((1, 8 * 8 * 4, 4)), MATS, 0, 0.0)
    assert "morphology holds" in str(e.value)


def test_batch_rejects_mismatched_lists():
    sim = _small()
    vox, out = _dev_bufs(sim)
    with pytest.raises(capi.CyrError):
        capi.simulate_batch([sim, sim], [vox], [MATS, MATS], [out, out])
    with pytest.raises(capi.CyrError) as e:
        capi.simulate_batch([sim], [vox], [MATS], [out])
    assert "not on a GPU" in str(e.value)


def test_noalloc_rejects_beyond_scratch_without_a_device():
    """The size gate runs on the host before any device call."""
    sim = capi.Sim((512, 512, 128), 4, 2.0, [285.0], (0.0, 30.0, 150.0), pipeline=capi.PIPE_MBASIS)
    buf = np.zeros(64)
    p = ctypes.c_void_p(buf.ctypes.data)
    rc = capi.lib().cyr_simulate_device_noalloc(sim.ref, p, p, p, None)
    assert rc == 1
    assert b"scratch" in capi.lib().cyr_last_error()
