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
