"""The allocation-free entry point (cyr_simulate_device_noalloc) against the device
API it mirrors (cyr_simulate_device), through the C ABI: the same executors on the
same inputs give bit-identical patterns.  Two calls in a row share the static
scratch array, the second with a smaller grid, so stale contents of the first
call must not leak into the second (a third, with voxels small enough that the
material basis skips detector columns outside the Ewald sphere, follows); a
simulation beyond the array's size is
rejected before any device work and leaves the output untouched."""
import numpy as np
import pytest

from cyrsoxs_amd import capi, synthetic
from tests.helpers import pattern_rel_l2, random_constants, random_euler

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _pair(sim, vox, mats, dtype):
    dv = torch.from_numpy(np.ascontiguousarray(vox, dtype=np.float64 if dtype == torch.float64 else np.float32)).cuda()
    nx, ny, _ = sim.dims
    shape = (len(sim.energies), sim.num_k, ny, nx)
    ref = torch.empty(shape, dtype=dtype, device="cuda")
    got = torch.full(shape, -7.0, dtype=dtype, device="cuda")
    capi.simulate_device(sim, dv, mats, ref)
    capi.simulate_device_noalloc(sim, dv, mats, got)
    torch.cuda.synchronize()
    return got.cpu().numpy(), ref.cpu().numpy()


@pytest.mark.parametrize("pipeline,prec", [(capi.PIPE_MBASIS, capi.FP32), (capi.PIPE_MBASIS, capi.FP64),
                                           (capi.PIPE_EBASIS, capi.FP32), (capi.PIPE_PLAIN, capi.FP32)])
def test_noalloc_matches_device_api_twice(pipeline, prec):
    dt = torch.float64 if prec == capi.FP64 else torch.float32
    # call 1: the 64^3 sphere, 3 energies x 6 angles
    c = synthetic.config0()
    print(f"noalloc pipeline={pipeline} prec={prec}: 64^3 sphere", flush=True)
    sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], interpolation=0, precision=prec,
                   pipeline=pipeline)
    got, ref = _pair(sim, c["voxels"], c["mats"], dt)
    assert np.array_equal(got, ref, equal_nan=True)
    # call 2: a smaller grid with other options (linear, Hanning, rotation mask)
    dims, M = (32, 24, 16), 3
    print(f"noalloc pipeline={pipeline} prec={prec}: 32x24x16", flush=True)
    vox = random_euler(dims, M, 91)
    mats = random_constants(2, M, 92)
    sim2 = capi.Sim(dims, M, 5.0, [283.0, 286.5], (0.0, 20.0, 160.0), interpolation=1, windowing=1, rot_mask=1,
                    precision=prec, pipeline=pipeline)
    got2, ref2 = _pair(sim2, vox, mats, dt)
    assert np.array_equal(got2, ref2, equal_nan=True)
    err, same = pattern_rel_l2(got2, ref2)
    assert same and err == 0.0
    # call 3: 1.6 nm voxels, so the detector corners fall outside the Ewald sphere
    # (the material basis's row spans then skip columns; Nz = 64 takes its Z band pass)
    dims3 = (48, 48, 64)
    print(f"noalloc pipeline={pipeline} prec={prec}: 48x48x64, 1.6 nm", flush=True)
    vox3 = random_euler(dims3, 1, 93)
    mats3 = random_constants(2, 1, 94)
    sim3 = capi.Sim(dims3, 1, 1.6, [283.0, 286.5], (0.0, 20.0, 160.0), interpolation=1, precision=prec,
                    pipeline=pipeline)
    got3, ref3 = _pair(sim3, vox3, mats3, dt)
    assert np.array_equal(got3, ref3, equal_nan=True)


def test_noalloc_rejects_beyond_scratch():
    dims = (512, 512, 128)
    sim = capi.Sim(dims, 4, 2.0, [285.0], (0.0, 30.0, 150.0), interpolation=1, pipeline=capi.PIPE_MBASIS)
    vox = torch.zeros((4, 512 * 512 * 128, 4), dtype=torch.float32, device="cuda")
    out = torch.full((1, 1, 512, 512), 3.0, dtype=torch.float32, device="cuda")
    with pytest.raises(capi.CyrError) as e:
        capi.simulate_device_noalloc(sim, vox, [[[1, 0, 1, 0]] * 4], out)
    assert e.value.status == 1 and "scratch" in str(e.value)
    torch.cuda.synchronize()
    assert bool((out == 3.0).all())
