"""The benchmark workloads' own generators and production kernel shapes against
the FP64 oracle (reference loop: src/cudaMain.cu:586-971).

* Bulk morphology (4 materials, Gaussian-filtered softmax volume fractions,
  filtered S, sphere-uniform Euler angles, linear Ewald, Hanning, 0:5:175):
  at 64^2 x 32 with every executor, and at the full 512^2 x 128 grid (two
  energies, angles cut to 0:30:150, which still spans two quarter turns) so the
  512-point X/Y and 128-point Z-band kernels and the quarter-turn orbit
  rotation are held to the oracle at their production shapes.
* A request served by the engine next to others (bulk morphology, AUTO
  executor) against the oracle.
* Grazing incidence (case type 2): tilted k vector and a detector direction
  away from z, so the detector warp is a genuine affine map
  (src/cudaMain.cu:916-960, include/Rotation.h:402-413).

Bars per pattern over the detector mask: FP64 <= 1e-10, FP32 <= 1e-4 relative
L2; NaN masks identical.
"""
import numpy as np
import pytest

from cyrsoxs_amd import capi, synthetic
from oracle import oracle as O
from tests.helpers import pattern_rel_l2, random_constants, random_euler

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

_CACHE = {}


def _check(got, ref, prec):
    err, same = pattern_rel_l2(got, ref)
    tol = 1e-10 if prec == capi.FP64 else 1e-4
    assert same, "NaN mask differs from the oracle"
    assert err < tol, f"relative L2 {err:.3e} > {tol}"
    return err


def _bulk(dims, M, seed, energies):
    vox = synthetic.bulk_morphology(dims, M, seed=seed, sigma=6.0)
    mats = synthetic.lorentzian_constants(energies, M, seed=seed)
    return vox, mats


@pytest.mark.parametrize("pipeline", [capi.PIPE_PLAIN, capi.PIPE_EBASIS, capi.PIPE_MBASIS, capi.PIPE_AUTO])
@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_bulk_reduced_grid(prec, pipeline):
    """Bulk workload generator at 64^2 x 32, 4 materials, 2 energies, full 0:5:175."""
    c = synthetic.config1_desc()
    dims, M = (64, 64, 32), 4
    energies = c["energies"][:2]
    print(f"bulk reduced prec={prec} pipeline={pipeline}", flush=True)
    if "bulk_small" not in _CACHE:
        vox, mats = _bulk(dims, M, 1, energies)
        vox64 = vox.numpy().astype(np.float64)
        osim = O.Sim(dims, M, c["phys_size"], energies, c["angles"], interpolation=1, windowing=1)
        _CACHE["bulk_small"] = (vox, mats, O.simulate(osim, vox64, mats))
    vox, mats, ref = _CACHE["bulk_small"]
    sim = capi.Sim(dims, M, c["phys_size"], energies, c["angles"], interpolation=1, windowing=1, precision=prec,
                   pipeline=pipeline)
    err = _check(capi.simulate(sim, vox.numpy(), mats), ref, prec)
    print(f"  rel L2 = {err:.3e}")


def test_bulk_full_grid():
    """512^2 x 128, 4 materials, PhysSize 2 nm, linear, Hanning: the headline's grid
    and kernel instantiations (FP32 material-basis and E-basis executors, FP64
    material basis), 280 and 281 eV, EAngle 0:30:150."""
    c = synthetic.config1_desc()
    dims, M = c["dims"], c["num_materials"]
    energies = c["energies"][:2]
    angles = (0.0, 30.0, 150.0)
    print("bulk full grid: generating", flush=True)
    vox = synthetic.bulk_morphology(dims, M, seed=1, sigma=6.0, device="cuda")
    mats = synthetic.lorentzian_constants(c["energies"], M, seed=1)[:2]
    vox64 = vox.double().cpu().numpy()
    print("bulk full grid: oracle", flush=True)
    osim = O.Sim(dims, M, c["phys_size"], energies, angles, interpolation=1, windowing=1)
    ref = O.simulate(osim, vox64, mats)
    del vox64
    out = torch.empty((2, 1, dims[1], dims[0]), dtype=torch.float32, device="cuda")
    for pipeline in (capi.PIPE_MBASIS, capi.PIPE_EBASIS):
        sim = capi.Sim(dims, M, c["phys_size"], energies, angles, interpolation=1, windowing=1, precision=capi.FP32,
                       pipeline=pipeline)
        print(f"bulk full grid: FP32 pipeline={pipeline}", flush=True)
        capi.simulate_device(sim, vox, mats, out)
        err = _check(out.cpu().numpy(), ref, capi.FP32)
        print(f"  rel L2 = {err:.3e}")
    print("bulk full grid: FP64 material basis", flush=True)
    vd = vox.double()
    del vox
    out64 = torch.empty((2, 1, dims[1], dims[0]), dtype=torch.float64, device="cuda")
    sim = capi.Sim(dims, M, c["phys_size"], energies, angles, interpolation=1, windowing=1, precision=capi.FP64,
                   pipeline=capi.PIPE_MBASIS)
    capi.simulate_device(sim, vd, mats, out64)
    err = _check(out64.cpu().numpy(), ref, capi.FP64)
    print(f"  rel L2 = {err:.3e}")


def test_served_request_matches_oracle():
    """Bulk-morphology requests served together (3 slots, AUTO executor): each one
    against the oracle."""
    reqs = []
    for i, (dims, M, nE) in enumerate([((64, 64, 32), 3, 3), ((48, 48, 16), 2, 5), ((64, 64, 16), 4, 2)]):
        energies = list(np.linspace(280.0, 305.0, nE))
        vox, mats = _bulk(dims, M, 100 + i, energies)
        sim = capi.Sim(dims, M, 2.0, energies, (0.0, 5.0, 175.0), interpolation=1, windowing=1,
                       pipeline=capi.PIPE_AUTO)
        reqs.append((sim, vox.numpy(), mats))
    with capi.Engine(slots=3) as eng:
        ids = [eng.submit(s, v, m) for s, v, m in reqs]
        for rid in ids:
            assert eng.wait(rid, timeout=60.0) == 0
        for rid, (s, v, m) in zip(ids, reqs):
            print(f"served {s.dims} resolved={s.resolved_pipeline()}", flush=True)
            osim = O.Sim(s.dims, s.desc.num_materials, 2.0, list(s.energies), (0.0, 5.0, 175.0), interpolation=1,
                         windowing=1)
            err = _check(eng.release(rid), O.simulate(osim, v.astype(np.float64), m), capi.FP32)
            print(f"  rel L2 = {err:.3e}")


@pytest.mark.parametrize("phys", [5.0, 1.5])
@pytest.mark.parametrize("pipeline", [capi.PIPE_PLAIN, capi.PIPE_EBASIS, capi.PIPE_MBASIS])
@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_grazing_incidence(prec, pipeline, phys):
    """Case type 2: k tilted off z and the detector normal tilted the other way (at
    1.5 nm voxels part of the detector lies outside the Ewald sphere)."""
    dims, M = (32, 32, 16), 2
    vox = random_euler(dims, M, 71)
    mats = random_constants(2, M, 72)
    k = np.array([[0.0, 0.12, 1.0]])
    k /= np.linalg.norm(k)
    det = np.array([0.15, -0.05, 1.0])
    det /= np.linalg.norm(det)
    kw = dict(interpolation=1, case_type=2)
    key = ("grazing", phys)
    if key not in _CACHE:
        osim = O.Sim(dims, M, phys, [284.0, 287.0], (0.0, 30.0, 150.0), kvectors=k, detector=det, **kw)
        _CACHE[key] = O.simulate(osim, vox, mats)
    ref = _CACHE[key]
    # the detector warp moves the pattern: it differs from the untilted detector's
    flat = O.simulate(O.Sim(dims, M, phys, [284.0, 287.0], (0.0, 30.0, 150.0), kvectors=k, **kw), vox, mats)
    assert not np.array_equal(np.isnan(flat), np.isnan(ref))
    print(f"grazing prec={prec} pipeline={pipeline} phys={phys}", flush=True)
    sim = capi.Sim(dims, M, phys, [284.0, 287.0], (0.0, 30.0, 150.0), kvectors=k, detector=det, precision=prec,
                   pipeline=pipeline, **kw)
    err = _check(capi.simulate(sim, vox, mats), ref, prec)
    print(f"  rel L2 = {err:.3e}")
