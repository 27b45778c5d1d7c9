"""Whole-pipeline parity of the HIP library against the FP64 oracle.

Bars (per pattern, over the detector mask = pixels finite in the oracle):
  GPU FP64  <= 1e-10 relative L2
  GPU FP32  <= 1e-4  relative L2
  NaN masks bit-identical in every precision and pipeline.
"""
import numpy as np
import pytest

from cyrsoxs_amd import capi, synthetic
from oracle import oracle as O
from tests.helpers import pattern_rel_l2, random_constants, random_euler, random_vector

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

_CACHE = {}


def _oracle(key, osim, vox, mats):
    if key not in _CACHE:
        _CACHE[key] = O.simulate(osim, vox, mats)
    return _CACHE[key]


def _check(got, ref, prec):
    err, same = pattern_rel_l2(got, ref)
    tol = 1e-10 if prec == capi.FP64 else 1e-4
    assert same, "NaN mask differs from the oracle"
    assert err < tol, f"relative L2 {err:.3e} > {tol}"
    return err


PIPES = [capi.PIPE_PLAIN, capi.PIPE_EBASIS, capi.PIPE_MBASIS]


@pytest.mark.parametrize("pipeline", PIPES + [capi.PIPE_AUTO])
@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_sphere_64(prec, pipeline):
    """The 64^3 two-material sphere, 3 energies x 6 angles, nearest neighbour."""
    c = synthetic.config0()
    print(f"sphere64 prec={prec} pipeline={pipeline}", flush=True)
    kw = dict(interpolation=c["interpolation"], windowing=c["windowing"])
    osim = O.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], **kw)
    ref = _oracle("sphere", osim, c["voxels"], c["mats"])
    sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], precision=prec, pipeline=pipeline, **kw)
    got = capi.simulate(sim, c["voxels"], c["mats"])
    err = _check(got, ref, prec)
    print(f"  rel L2 = {err:.3e}")


CASES = {
    "linear_hann": dict(dims=(32, 32, 16), M=3, morph=0, kw=dict(interpolation=1, windowing=1)),
    "rotmask_material": dict(dims=(32, 32, 16), M=2, morph=0, kw=dict(interpolation=1, rot_mask=1,
                                                                        reference_frame=0)),
    "vector_nn": dict(dims=(24, 24, 12), M=2, morph=1, kw=dict(interpolation=0)),
    "alg1_full": dict(dims=(32, 32, 8), M=3, morph=0, kw=dict(interpolation=1, algorithm=1, scatter_approach=1,
                                                                windowing=1)),
    "thin2d": dict(dims=(64, 64, 1), M=3, morph=0, kw=dict(interpolation=1)),
    "thin2d_rotmask": dict(dims=(48, 48, 1), M=2, morph=1, kw=dict(interpolation=0, rot_mask=1)),
    "odd_grid": dict(dims=(30, 30, 10), M=2, morph=0, kw=dict(interpolation=1)),
    "radix3_384": dict(dims=(384, 384, 4), M=2, morph=0, kw=dict(interpolation=1, windowing=1)),
    "pow2_128": dict(dims=(128, 128, 6), M=2, morph=1, kw=dict(interpolation=0)),
    "one_material": dict(dims=(32, 32, 16), M=1, morph=0, kw=dict(interpolation=1, windowing=1)),
    "five_materials": dict(dims=(24, 24, 16), M=5, morph=1, kw=dict(interpolation=1, reference_frame=0)),
    "deep_z": dict(dims=(32, 24, 128), M=3, morph=0, kw=dict(interpolation=1, windowing=1)),
    "deep_z_nn": dict(dims=(24, 32, 64), M=2, morph=1, kw=dict(interpolation=0)),
    # run-time X plan (no compile-time codelet for 288) with rows wider than a block
    "rt_wide_rows": dict(dims=(288, 288, 4), M=3, morph=0, kw=dict(interpolation=1, windowing=1)),
    # 720 angles: square (quarter-turn orbits, 180 steps each) and not square
    "many_angles": dict(dims=(24, 24, 8), M=2, morph=0, kw=dict(interpolation=1, windowing=1),
                        angles=(0.0, 0.5, 359.5)),
    "many_angles_rect": dict(dims=(40, 24, 1), M=2, morph=1, kw=dict(interpolation=1, rot_mask=1),
                             angles=(0.0, 0.5, 359.5)),
    # voxels small enough that the detector's corners lie outside the Ewald sphere
    # (|q| up to pi / PhysSize above |k| ~ 1.44 nm^-1): the material-basis Y pass
    # writes, and its Z band pass reads, only the columns some band reads; rows
    # and whole 32-column tiles with none are skipped
    "span_partial": dict(dims=(96, 64, 64), M=3, morph=0, kw=dict(interpolation=1, windowing=1), phys=1.2),
    # Nx = 80: the last 32-column tile of each row is ragged
    "span_partial_nn": dict(dims=(80, 64, 128), M=2, morph=1, kw=dict(interpolation=0), phys=1.6),
    # Ny = 26 has no compile-time plan (no span in the Y pass), Nz = 64 has
    "span_zband_only": dict(dims=(30, 26, 64), M=2, morph=0, kw=dict(interpolation=1), phys=1.3),
    # Nz = 32: no Z band pass (Z FFT pass + band extraction behind the spanned Y pass)
    "span_extract": dict(dims=(64, 64, 32), M=2, morph=0, kw=dict(interpolation=1), phys=1.4),
    "span_material_frame": dict(dims=(64, 48, 64), M=2, morph=0, kw=dict(interpolation=1, reference_frame=0),
                                phys=1.5),
    # material frame with the rotation mask on a 2D film at 1 nm (most of the detector outside
    # the Ewald sphere: the per-angle projections' rotation fills the tiles beyond it)
    "span_material_rotmask": dict(dims=(48, 48, 1), M=2, morph=1,
                                  kw=dict(interpolation=1, rot_mask=1, reference_frame=0)),
    # odd Nx (run-time X plan; the bin -> shifted map of odd lengths; a ragged last tile)
    "span_odd_nx": dict(dims=(45, 64, 64), M=2, morph=1, kw=dict(interpolation=1, windowing=1), phys=1.3),
}


@pytest.mark.parametrize("pipeline", PIPES)
@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("case", sorted(CASES))
def test_modes(case, prec, pipeline):
    cfg = CASES[case]
    print(f"mode {case} prec={prec} pipeline={pipeline}", flush=True)
    dims, M = cfg["dims"], cfg["M"]
    vox = random_euler(dims, M, 21) if cfg["morph"] == 0 else random_vector(dims, M, 21)
    energies = [282.0, 285.5]
    mats = random_constants(len(energies), M, 4)
    angles = cfg.get("angles", (0.0, 20.0, 160.0))
    kw = dict(cfg["kw"], morphology_type=cfg["morph"])
    phys = cfg.get("phys", 1.0 if dims[2] == 1 else 5.0)
    osim = O.Sim(dims, M, phys, energies, angles, **kw)
    ref = _oracle(case, osim, vox, mats)
    sim = capi.Sim(dims, M, phys, energies, angles, precision=prec, pipeline=pipeline, **kw)
    got = capi.simulate(sim, vox, mats)
    err = _check(got, ref, prec)
    print(f"  rel L2 = {err:.3e}")


# Quarter-turn orbit rotation (laboratory frame of the material-basis executor):
# angle lists spanning 2, 3 and 4 quarter turns, with exact multiples of 90
# degrees (back-mapped points on pixel edges, where a pair gathers its own taps)
# and without, rotation mask on and off.
ORBIT_CASES = {
    "q2_linear_hann": dict(dims=(48, 48, 8), angles=(0.0, 15.0, 165.0), kw=dict(interpolation=1, windowing=1)),
    "q3_nn_rotmask": dict(dims=(40, 40, 8), angles=(0.0, 30.0, 240.0), kw=dict(interpolation=0, rot_mask=1)),
    "q4_thin_linear": dict(dims=(64, 64, 1), angles=(0.0, 10.0, 350.0), kw=dict(interpolation=1)),
    "q4_offset_rotmask": dict(dims=(32, 32, 8), angles=(7.0, 18.0, 349.0), kw=dict(interpolation=1, rot_mask=1)),
    # more than 64 angles per quarter turn (the exact-path record spans several words)
    "q4_half_deg": dict(dims=(32, 32, 1), angles=(0.0, 0.5, 359.5), kw=dict(interpolation=1)),
    "q4_fine_rotmask": dict(dims=(24, 24, 4), angles=(0.0, 1.25, 358.75), kw=dict(interpolation=1, rot_mask=1)),
}


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("case", sorted(ORBIT_CASES))
def test_quarter_turn_orbits(case, prec):
    c = ORBIT_CASES[case]
    print(f"orbit {case} prec={prec}", flush=True)
    dims, M = c["dims"], 2
    vox = random_euler(dims, M, 31)
    energies = [283.0, 286.0, 289.0]
    mats = random_constants(len(energies), M, 5)
    phys = 1.0 if dims[2] == 1 else 5.0
    ref = _oracle("orbit_" + case, O.Sim(dims, M, phys, energies, c["angles"], **c["kw"]), vox, mats)
    sim = capi.Sim(dims, M, phys, energies, c["angles"], precision=prec, pipeline=capi.PIPE_MBASIS, **c["kw"])
    err = _check(capi.simulate(sim, vox, mats), ref, prec)
    print(f"  rel L2 = {err:.3e}")


def test_host_upload_matches_device_api():
    """cyr_simulate from a pageable host morphology above the staged-upload size
    (256^2 x 128, 2 materials: 268 MB, copied through pinned buffers by several
    host threads) returns exactly the device API's patterns; FP16 storage (the
    host converts, 134 MB) matches the FP32 run to its storage error."""
    dims, M = (256, 256, 128), 2
    print("host upload 256^2 x 128", flush=True)
    vox = synthetic.bulk_morphology(dims, M, seed=4, sigma=4.0)
    energies = [284.0, 286.0, 288.0]
    mats = synthetic.lorentzian_constants(energies, M, seed=4)
    kw = dict(interpolation=1, windowing=1, pipeline=capi.PIPE_MBASIS)
    sim = capi.Sim(dims, M, 5.0, energies, (0.0, 15.0, 165.0), **kw)
    host = capi.simulate(sim, vox.numpy(), mats)
    d_out = torch.empty((len(energies), 1, dims[1], dims[0]), dtype=torch.float32, device="cuda")
    capi.simulate_device(sim, vox.cuda(), mats, d_out)
    dev = d_out.cpu().numpy()
    assert np.array_equal(np.isnan(host), np.isnan(dev))
    fin = np.isfinite(dev)
    assert np.array_equal(host[fin], dev[fin])
    half = capi.simulate(capi.Sim(dims, M, 5.0, energies, (0.0, 15.0, 165.0), storage=capi.STORE_FP16, **kw),
                         vox.numpy(), mats)
    err, same = pattern_rel_l2(half, dev)
    print(f"host upload FP16 storage vs FP32: rel L2 {err:.3e}")
    assert same and err < 2e-3, err


@pytest.mark.parametrize("phys", [5.0, 1.4])
@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_beam_divergence_kvectors(prec, phys):
    """Case type 1: two tilted k vectors (the reference marks it experimental);
    at 1.4 nm voxels the Ewald sphere leaves an off-centre part of each detector
    row unread."""
    dims, M = (24, 24, 24), 2
    vox = random_euler(dims, M, 8)
    mats = random_constants(1, M, 9)
    k = np.array([[0.0, 0.1, 1.0], [0.05, 0.0, 1.0]])
    k = k / np.linalg.norm(k, axis=1, keepdims=True)
    kw = dict(interpolation=1, case_type=1)
    osim = O.Sim(dims, M, phys, [285.0], (0.0, 45.0, 135.0), kvectors=k, **kw)
    ref = O.simulate(osim, vox, mats)
    for pipeline in PIPES:
        sim = capi.Sim(dims, M, phys, [285.0], (0.0, 45.0, 135.0), kvectors=k, precision=prec, pipeline=pipeline,
                       **kw)
        got = capi.simulate(sim, vox, mats)
        _check(got, ref, prec)


def test_device_api_matches_host_api():
    c = synthetic.config0()
    sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], interpolation=0,
                   pipeline=capi.PIPE_EBASIS)
    host = capi.simulate(sim, c["voxels"], c["mats"])
    dv = torch.from_numpy(c["voxels"].astype(np.float32)).cuda()
    out = torch.empty((3, 1, 64, 64), dtype=torch.float32, device="cuda")
    capi.simulate_device(sim, dv, c["mats"], out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), host)


def test_repeatable_bitwise():
    """No atomics in any accumulation: two runs are bit-identical."""
    c = synthetic.config0()
    for pipeline in PIPES:
        sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], interpolation=0,
                       pipeline=pipeline)
        a = capi.simulate(sim, c["voxels"], c["mats"])
        b = capi.simulate(sim, c["voxels"], c["mats"])
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("storage,tol", [(capi.STORE_FP16, 1e-4), (capi.STORE_BF16, 4e-4)])
def test_half_precision_morphology_storage(storage, tol):
    """FP16/BF16 morphology storage (FP32 compute): masks identical; bars about 10x the
    measured errors (FP16 9.3e-6, BF16 3.3e-5 relative L2 on MI355X)."""
    dims, M = (32, 32, 16), 3
    vox = random_euler(dims, M, 31)
    energies = [284.0, 286.0]
    mats = random_constants(2, M, 5)
    kw = dict(interpolation=1, windowing=1)
    ref = _oracle("half_storage", O.Sim(dims, M, 5.0, energies, (0, 30, 150), **kw), vox, mats)
    for pipeline in PIPES:
        sim = capi.Sim(dims, M, 5.0, energies, (0, 30, 150), precision=capi.FP32, storage=storage, pipeline=pipeline,
                       **kw)
        got = capi.simulate(sim, vox, mats)
        err, same = pattern_rel_l2(got, ref)
        print(f"storage={storage} pipeline={pipeline}: rel L2 {err:.3e}")
        assert same and err < tol


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_mbasis_energy_groups(prec):
    """Energies spread far enough that their Ewald bands exceed the per-group cap:
    the material-basis executor splits them into groups (one basis transform each)."""
    dims, M = (32, 32, 32), 4
    vox = random_euler(dims, M, 41)
    energies = [200.0, 240.0, 290.0, 350.0, 420.0, 500.0]
    mats = random_constants(len(energies), M, 6)
    kw = dict(interpolation=1, windowing=1)
    ref = _oracle("energy_groups", O.Sim(dims, M, 2.0, energies, (0.0, 40.0, 160.0), **kw), vox, mats)
    sim = capi.Sim(dims, M, 2.0, energies, (0.0, 40.0, 160.0), precision=prec, pipeline=capi.PIPE_MBASIS, **kw)
    capi.kernel_timing(True)
    try:
        got = capi.simulate(sim, vox, mats)
        launches = {k: v["calls"] for k, v in capi.kernel_timing_report().items()}
    finally:
        capi.kernel_timing(False)
    err = _check(got, ref, prec)
    print(f"energy groups prec={prec}: rel L2 {err:.3e}, launches {launches}")
    # more than one group: more band passes than the one a single group needs
    assert launches.get("mbasis_bands", 0) > 1


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_mbasis_matches_ebasis(prec):
    """The two fused executors agree (same pipeline, linear steps in another order)."""
    c = synthetic.config0()
    kw = dict(interpolation=1, windowing=1)
    out = {}
    for pipeline in (capi.PIPE_EBASIS, capi.PIPE_MBASIS):
        sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], precision=prec, pipeline=pipeline,
                       **kw)
        out[pipeline] = capi.simulate(sim, c["voxels"], c["mats"])
    err, same = pattern_rel_l2(out[capi.PIPE_MBASIS], out[capi.PIPE_EBASIS])
    assert same
    assert err < (1e-11 if prec == capi.FP64 else 2e-5), err


def test_mbasis_large_plane_rotation():
    """A 1280 x 1152 film (not square, so no quarter-turn orbits): quadratic-form
    planes above 16 MB take the one-energy-per-pass rotation; held against the
    E-basis executor's per-angle rotation."""
    nx, ny, M = 1280, 1152, 2
    assert nx * ny * 12 > 16 << 20
    vox = random_euler((nx, ny, 1), M, 61)
    energies = [284.0, 286.0, 288.5]
    mats = random_constants(len(energies), M, 62)
    out = {}
    for pipeline in (capi.PIPE_EBASIS, capi.PIPE_MBASIS):
        sim = capi.Sim((nx, ny, 1), M, 1.0, energies, (0.0, 5.0, 175.0), interpolation=1, precision=capi.FP32,
                       pipeline=pipeline)
        out[pipeline] = capi.simulate(sim, vox, mats)
    err, same = pattern_rel_l2(out[capi.PIPE_MBASIS], out[capi.PIPE_EBASIS])
    print(f"large plane: rel L2 {err:.3e}")
    assert same and err < 1e-4  # two FP32 paths, each held to 1e-4 of the oracle


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32])
def test_mbasis_nz256(prec):
    """Nz = 256 pencils in the Z band pass (FP32: 32-column tiles of 1024 threads),
    with a row width that leaves part of each tile's columns idle."""
    dims, M = (48, 40, 256), 2
    vox = random_euler(dims, M, 71)
    energies = [283.0, 286.5]
    mats = random_constants(len(energies), M, 72)
    kw = dict(interpolation=1, windowing=1)
    ref = _oracle("nz256", O.Sim(dims, M, 2.0, energies, (0.0, 30.0, 150.0), **kw), vox, mats)
    sim = capi.Sim(dims, M, 2.0, energies, (0.0, 30.0, 150.0), precision=prec, pipeline=capi.PIPE_MBASIS, **kw)
    err = _check(capi.simulate(sim, vox, mats), ref, prec)
    print(f"nz256 prec={prec}: rel L2 {err:.3e}")


@pytest.mark.parametrize("phys", [5.0, 1.6])
def test_mbasis_row_spans(phys):
    """Columns whose detector pixels lie outside every energy's Ewald sphere are
    neither written by the Y pass nor read by the Z band pass: with 5 nm voxels
    (|q| <= pi/5 nm^-1 < |k|) every column is kept, with 1.6 nm voxels the corners
    drop out (about 43% of the columns kept at 48 x 64).  The patterns match the
    oracle either way."""
    dims, M = (48, 64, 64), 2  # (Nz = 64: the fused Z band pass)
    vox = random_euler(dims, M, 81)
    energies = [282.0, 285.5]
    mats = random_constants(len(energies), M, 82)
    kw = dict(interpolation=1, windowing=1)
    ref = _oracle(f"spans_{phys}", O.Sim(dims, M, phys, energies, (0.0, 20.0, 160.0), **kw), vox, mats)
    sim = capi.Sim(dims, M, phys, energies, (0.0, 20.0, 160.0), precision=capi.FP32, pipeline=capi.PIPE_MBASIS,
                   **kw)
    capi.kernel_timing(True)
    try:
        got = capi.simulate(sim, vox, mats)
        rep = capi.kernel_timing_report()
    finally:
        capi.kernel_timing(False)
    _check(got, ref, capi.FP32)
    reads = 7.0 * np.prod(dims) * 8  # one material pair: 7 packed complex64 volumes
    kept = rep["fft_y"]["avg_bytes"] / reads - 1.0  # written share of the columns
    print(f"phys {phys}: Y pass keeps {kept:.3f} of the columns; Z band bytes {rep['mbasis_zband']['avg_bytes']:.0f}")
    # the share of detector pixels inside the Ewald circle of the highest energy
    # (beam along z: |q_perp| <= |k|, hc = 1239.84 eV nm)
    kmax = 2 * np.pi * max(energies) / 1239.84193
    qx = (np.arange(dims[0]) - dims[0] // 2) * 2 * np.pi / (dims[0] * phys)
    qy = (np.arange(dims[1]) - dims[1] // 2) * 2 * np.pi / (dims[1] * phys)
    disc = float(np.mean(qx[None, :] ** 2 + qy[:, None] ** 2 <= kmax**2))
    assert abs(kept - disc) < 0.01, (kept, disc)
    if phys == 5.0:
        assert kept > 0.995  # (a detector corner can still fall outside the sampled bins)
    else:
        assert 0.3 < kept < 0.6
        # the Z band pass reads only those columns (plus its band-value writes)
        assert rep["mbasis_zband"]["avg_bytes"] < 0.7 * reads


def test_energy_lanes_bitwise():
    """One or two (energy, k) units in flight give bit-identical patterns."""
    c = synthetic.config0()
    sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], interpolation=1,
                   pipeline=capi.PIPE_EBASIS)
    lanes = capi.energy_lanes()
    try:
        capi.set_energy_lanes(1)
        a = capi.simulate(sim, c["voxels"], c["mats"])
        capi.set_energy_lanes(2)
        b = capi.simulate(sim, c["voxels"], c["mats"])
    finally:
        capi.set_energy_lanes(lanes)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("columns,target", [(4, 256), (8, 2048), (16, 512)])
def test_yz_tuning_changes_rounding_only(columns, target):
    """Every Y/z launch shape the autotuner may pick meets the FP32 bar."""
    c = synthetic.config0()
    kw = dict(interpolation=1, windowing=0)
    osim = O.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], **kw)
    ref = _oracle("sphere_linear", osim, c["voxels"], c["mats"])
    sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], pipeline=capi.PIPE_EBASIS, **kw)
    try:
        capi.set_tuning(c["dims"], capi.FP32, columns, target)
        got = capi.simulate(sim, c["voxels"], c["mats"])
    finally:
        capi.set_tuning(c["dims"], capi.FP32, 0, 0)
    _check(got, ref, capi.FP32)


def test_autotune_records_a_choice():
    c = synthetic.config0()
    sim = capi.Sim(c["dims"], 2, c["phys_size"], c["energies"], c["angles"], interpolation=1,
                   pipeline=capi.PIPE_EBASIS)
    try:
        cols, target = capi.autotune(sim)
        assert cols in (4, 8, 16) and target > 0
        assert capi.get_tuning(c["dims"], capi.FP32) == (cols, target)
    finally:
        capi.set_tuning(c["dims"], capi.FP32, 0, 0)


@pytest.mark.parametrize("prec,storage,tol", [(capi.FP64, capi.STORE_NATIVE, 1e-10), (capi.FP32, capi.STORE_NATIVE, 1e-4),
                                              (capi.FP32, capi.STORE_FP16, 3e-4), (capi.FP32, capi.STORE_BF16, 3e-3)])
def test_thinfilm_precision(prec, storage, tol):
    """The thin-film morphology (3 phases, in-plane director) at 256^2 x 1 with the
    full 0:2:358 rotation against the FP64 oracle; FP16/BF16 morphology storage
    carries its own error (measured 2.1e-5 and 2.5e-4 relative L2; bars about 10x)."""
    n = 256
    vox = synthetic.thinfilm_morphology(n, seed=3).numpy().astype(np.float64)
    energies = [284.0, 286.5]
    mats = synthetic.lorentzian_constants(energies, 3, seed=3)
    angles = (0.0, 2.0, 358.0)
    ref = _oracle("thinfilm", O.Sim((n, n, 1), 3, 1.0, energies, angles, interpolation=1), vox, mats)
    for pipeline in (capi.PIPE_EBASIS, capi.PIPE_MBASIS):
        sim = capi.Sim((n, n, 1), 3, 1.0, energies, angles, interpolation=1, precision=prec, storage=storage,
                       pipeline=pipeline)
        got = capi.simulate(sim, vox, mats)
        err, same = pattern_rel_l2(got, ref)
        print(f"thin film prec={prec} storage={storage} pipeline={pipeline}: rel L2 {err:.3e}")
        assert same and err < tol
