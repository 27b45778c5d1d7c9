"""Batched execution, the graph-captured batch plan and the serving engine on the
GPU.  Bars: every member of a batch, and every request served in a mix, matches
the same simulation run alone to 1e-4 relative L2 per pattern with identical
NaN masks (the FP32 bar against the oracle); graph replays are bit-identical to
the eager run of the same plan."""
import numpy as np
import pytest

from cyrsoxs_amd import capi
from oracle import oracle as O
from tests.helpers import pattern_rel_l2, random_constants, random_euler, random_vector

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MEMBERS = [
    dict(dims=(32, 32, 16), M=2, morph=0, pipeline=capi.PIPE_EBASIS, kw=dict(interpolation=1, windowing=1)),
    dict(dims=(48, 48, 24), M=3, morph=0, pipeline=capi.PIPE_EBASIS, kw=dict(interpolation=0)),
    dict(dims=(64, 64, 1), M=2, morph=1, pipeline=capi.PIPE_MBASIS, kw=dict(interpolation=1)),
    dict(dims=(24, 24, 12), M=2, morph=0, pipeline=capi.PIPE_PLAIN, kw=dict(interpolation=1, rot_mask=1)),
    dict(dims=(64, 64, 32), M=4, morph=0, pipeline=capi.PIPE_EBASIS, kw=dict(interpolation=1)),
    dict(dims=(40, 40, 64), M=3, morph=1, pipeline=capi.PIPE_MBASIS, kw=dict(interpolation=1, windowing=1)),
]


def _member(i, cfg, nE=None):
    dims, M = cfg["dims"], cfg["M"]
    vox = random_euler(dims, M, 40 + i) if cfg["morph"] == 0 else random_vector(dims, M, 40 + i)
    energies = [283.0 + 0.7 * j for j in range(nE or (2 + i % 3))]
    mats = random_constants(len(energies), M, 50 + i)
    phys = 1.0 if dims[2] == 1 else 4.0
    sim = capi.Sim(dims, M, phys, energies, (0.0, 15.0, 165.0), morphology_type=cfg["morph"],
                   pipeline=cfg["pipeline"], **cfg["kw"])
    return sim, vox.astype(np.float32), mats


def _alone(sim, vox, mats):
    return capi.simulate(sim, vox, mats)


def _same(got, ref, tol=1e-4):
    err, same = pattern_rel_l2(got, ref)
    assert same, "NaN mask differs"
    assert err < tol, f"relative L2 {err:.3e}"
    return err


def test_simulate_batch_matches_members_alone():
    members = [_member(i, c) for i, c in enumerate(MEMBERS)]
    d_vox = [torch.from_numpy(v).cuda() for _, v, _ in members]
    d_out = [torch.empty((len(s.energies), 1, s.dims[1], s.dims[0]), dtype=torch.float32, device="cuda")
             for s, _, _ in members]
    capi.simulate_batch([s for s, _, _ in members], d_vox, [m for _, _, m in members], d_out)
    torch.cuda.synchronize()
    for (s, v, m), o in zip(members, d_out):
        print(f"batch member {s.dims} pipeline={s.desc.pipeline}", flush=True)
        _same(o.cpu().numpy(), _alone(s, v, m))


def test_batch_member_matches_oracle():
    s, v, m = _member(0, MEMBERS[0])
    d_vox, d_out = torch.from_numpy(v).cuda(), torch.empty((len(s.energies), 1, 32, 32), device="cuda")
    capi.simulate_batch([s], [d_vox], [m], [d_out])
    osim = O.Sim(s.dims, 2, 4.0, list(s.energies), (0.0, 15.0, 165.0), **MEMBERS[0]["kw"])
    _same(d_out.cpu().numpy(), O.simulate(osim, v.astype(np.float64), m))


def test_batch_plan_graph_replays():
    """Every replay of the captured graph equals the eager result; a replay after
    the morphology changed in place equals a fresh run on the new data."""
    members = [_member(i, c) for i, c in enumerate(MEMBERS) if i in (0, 2, 4, 5)]
    d_vox = [torch.from_numpy(v).cuda() for _, v, _ in members]
    d_out = [torch.empty((len(s.energies), 1, s.dims[1], s.dims[0]), dtype=torch.float32, device="cuda")
             for s, _, _ in members]
    plan = capi.BatchPlan([s for s, _, _ in members], d_vox, [m for _, _, m in members], d_out)
    try:
        torch.cuda.synchronize()
        eager = [o.cpu().numpy().copy() for o in d_out]
        for (s, v, m), e in zip(members, eager):
            _same(e, _alone(s, v, m))
        for rep in range(3):
            for o in d_out:
                o.fill_(0.0)
            plan.launch()
            torch.cuda.synchronize()
            for o, e in zip(d_out, eager):
                np.testing.assert_array_equal(o.cpu().numpy(), e)
            # the program's own eager work between replays
            s0, v0, m0 = members[0]
            _alone(s0, v0, m0)
        # new data in the bound buffers
        s0, v0, m0 = members[0]
        v_new = v0.copy()
        v_new[[0, 1]] = v_new[[1, 0]]
        d_vox[0].copy_(torch.from_numpy(v_new))
        plan.launch()
        torch.cuda.synchronize()
        _same(d_out[0].cpu().numpy(), _alone(s0, v_new, m0))
    finally:
        plan.close()


def test_engine_results_alone_and_in_mix():
    members = [_member(i, c) for i, c in enumerate(MEMBERS)] + [_member(5, MEMBERS[4], nE=5)]
    alone = [_alone(s, v, m) for s, v, m in members]
    with capi.Engine(slots=3) as eng:
        ids = [eng.submit(s, v, m) for s, v, m in members]
        for rid in ids:
            assert eng.wait(rid, timeout=60.0) == 0
        for rid, ref in zip(ids, alone):
            q, r, lat = eng.times(rid)
            assert 0 <= q <= lat and 0 < r <= lat
            _same(eng.release(rid), ref)
        st = eng.stats()
        assert st["submitted"] == len(members) and st["completed"] == len(members)
        assert st["in_flight"] == 0 and st["queued"] == 0
        # idle slots keep their buffers, and they stay counted against the budget
        assert 0 < st["resident_bytes"] <= st["budget_bytes"]


def _blocker():
    """A plain-pipeline request that keeps one slot busy for a while (128^3, 4
    energies x 36 angles: 144 polarization + 3 FFT passes)."""
    sim, vox, mats = _member(7, dict(dims=(128, 128, 128), M=2, morph=0, pipeline=capi.PIPE_PLAIN,
                                     kw=dict(interpolation=1)), nE=4)
    sim = capi.Sim(sim.dims, 2, 4.0, list(sim.energies), (0.0, 5.0, 175.0), pipeline=capi.PIPE_PLAIN, interpolation=1)
    return sim, vox, mats


def test_engine_admission_cancel_and_timeout():
    s, v, m = _member(4, MEMBERS[4], nE=12)
    bs, bv, bm = _blocker()
    with capi.Engine(slots=1) as eng:
        with capi.Engine(slots=1, budget_bytes=s.footprint() + 1024) as small:
            big, vb, mb = _member(4, dict(MEMBERS[4], dims=(96, 96, 48)))
            with pytest.raises(capi.CyrError) as e:
                small.submit(big, vb, mb)
            assert "ADMISSION" in str(e.value)
            assert small.stats()["rejected"] == 1
        # the only slot runs the blocker; the next request waits in the queue
        blocker = eng.submit(bs, bv, bm)
        queued = eng.submit(s, v, m)
        last = eng.submit(s, v, m)
        with pytest.raises(capi.CyrError) as e:
            eng.wait(queued, timeout=0.0)
        assert e.value.status == 8  # CYR_E_TIMEOUT: still queued
        eng.cancel(last)
        assert eng.wait(last, timeout=60.0) == 5  # CYR_E_CANCELLED
        assert eng.wait(blocker, timeout=120.0) == 0
        assert eng.wait(queued, timeout=60.0) == 0
        _same(eng.result(queued), _alone(s, v, m))
        stats = eng.stats()
        assert stats["cancelled"] == 1 and stats["completed"] == 2 and stats["queued"] == 0


def test_engine_counts_retained_buffers():
    """A large request then small ones, with a budget that holds the large one's
    footprint only: the slot that ran it frees its buffers before a small one runs
    (otherwise the device would hold both), every request succeeds, and the
    reported reservation always covers what the idle slots hold."""
    big, vb, mb = _member(4, dict(MEMBERS[4], dims=(96, 96, 48)), nE=3)
    small = [_member(i, MEMBERS[i], nE=2) for i in (0, 1, 2)]
    held_big = vb.nbytes + 3 * 96 * 96 * 4  # its morphology and patterns at least
    with capi.Engine(slots=2, budget_bytes=big.footprint() + 1024) as eng:
        rid = eng.submit(big, vb, mb)
        assert eng.wait(rid, timeout=60.0) == 0
        st = eng.stats()
        assert held_big <= st["resident_bytes"] <= st["budget_bytes"]
        _same(eng.release(rid), _alone(big, vb, mb))
        ids = [eng.submit(sm, v, m) for sm, v, m in small]
        for (sm, v, m), i in zip(small, ids):
            assert eng.wait(i, timeout=60.0) == 0
            _same(eng.release(i), _alone(sm, v, m))
        st = eng.stats()
        assert 0 < st["resident_bytes"] <= st["budget_bytes"]
        assert st["completed"] == 4 and st["in_flight"] == 0


def test_release_workspace_refused_while_plan_or_engine_lives():
    members = [_member(0, MEMBERS[0])]
    d_vox = [torch.from_numpy(v).cuda() for _, v, _ in members]
    d_out = [torch.empty((len(s.energies), 1, s.dims[1], s.dims[0]), dtype=torch.float32, device="cuda")
             for s, _, _ in members]
    plan = capi.BatchPlan([s for s, _, _ in members], d_vox, [m for _, _, m in members], d_out)
    try:
        assert capi.lib().cyr_release_workspace() == 1  # CYR_E_INVALID: the graph holds cache pointers
        assert b"batch plan" in capi.lib().cyr_last_error()
    finally:
        plan.close()
    with capi.Engine(slots=1):
        assert capi.lib().cyr_release_workspace() == 1
    assert capi.lib().cyr_release_workspace() == 0
    # the library rebuilds its caches on the next call
    s, v, m = members[0]
    _same(_alone(s, v, m), O.simulate(O.Sim(s.dims, 2, 4.0, list(s.energies), (0.0, 15.0, 165.0),
                                            **MEMBERS[0]["kw"]), v.astype(np.float64), m))


def test_injected_allocation_failure_is_reported_and_recovered(monkeypatch):
    """A device allocation that fails (forced from the host through the allocator's
    test hook, CYR_TEST_FAIL_ALLOC) makes the call return CYR_E_NOMEM; the library
    and a serving engine keep working afterwards."""
    s, v, m = _member(1, MEMBERS[1])
    ref = O.simulate(O.Sim(s.dims, 3, 4.0, list(s.energies), (0.0, 15.0, 165.0), **MEMBERS[1]["kw"]),
                     v.astype(np.float64), m)
    assert capi.lib().cyr_release_workspace() == 0  # empty caches: the next call allocates
    monkeypatch.setenv("CYR_TEST_FAIL_ALLOC", "2")  # the third allocation fails
    with pytest.raises(capi.CyrError) as e:
        capi.simulate(s, v, m)
    assert e.value.status == 2 and "injected" in str(e.value)
    monkeypatch.delenv("CYR_TEST_FAIL_ALLOC")
    _same(capi.simulate(s, v, m), ref)
    # in the engine: the failing request reports CYR_E_NOMEM, the next one is served
    with capi.Engine(slots=1) as eng:
        monkeypatch.setenv("CYR_TEST_FAIL_ALLOC", "0")
        bad = eng.submit(s, v, m)
        assert eng.wait(bad, timeout=60.0) == 2
        monkeypatch.delenv("CYR_TEST_FAIL_ALLOC")
        good = eng.submit(s, v, m)
        assert eng.wait(good, timeout=60.0) == 0
        _same(eng.result(good), ref)
        assert eng.stats()["completed"] == 2
