"""Batched workload: 64 independent morphologies of uneven size in one batch.

Grids are drawn per seed from {128^2 x 32, 256^2 x 64, 384^2 x 48, 512^2 x 64},
each member with 2-4 materials (Gaussian-filtered Vfrac normalised to 1, random
S, sphere-uniform Euler angles), 8-32 energies over 280-300 eV and 18 E angles
(0:10:170), linear Ewald interpolation, no window, FP32.  Throughput is
voxel*angle*energy per second over the whole batch (inputs resident in HBM).

Modes, all on the same device buffers (P = the executor under test, --pipeline,
default the material-basis executor):
  plain_serial    the reference's per-angle structure, one simulation at a time
  ebasis_serial   the E-basis executor, one simulation at a time (its outputs are
                  the comparison reference of the other modes)
  P_serial        executor P, one simulation at a time
  P_batch         cyr_simulate_batch (members over 4 streams, shapes autotuned)
  P_graph         a cyr_batch plan (the same schedule captured in a hipGraph),
                  replayed

    python -m cyrsoxs_amd.bench_batch [--count 64] [--plain-count N] [--reps 2] [--pipeline mbasis|ebasis]

One JSON line on stdout.  --plain-count times the plain path on the first N
members (in the batch's own order) and scales by their share of the units when
N < count; the JSON says which.
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--plain-count", type=int, default=16)
    ap.add_argument("--pipeline", choices=("mbasis", "ebasis"), default="mbasis")
    return ap.parse_args(argv)


def main(argv=None):
    print(json.dumps(run(parse(argv))), flush=True)
    return 0


def run(a):
    """The batched workload's modes; returns the result object (the JSON line)."""
    import numpy as np
    import torch

    from cyrsoxs_amd import capi, synthetic

    torch.cuda.set_device(0)
    pn = a.pipeline
    PIPE = capi.PIPE_MBASIS if pn == "mbasis" else capi.PIPE_EBASIS
    members = synthetic.config3_members(a.count, a.seed)
    sims, vox, mats, outs = [], [], [], []
    units = []
    for m in members:
        nx, ny, nz = m["dims"]
        M = m["num_materials"]
        vox.append(synthetic.bulk_morphology(m["dims"], M, seed=m["seed"], sigma=6.0, device="cuda"))
        mats.append(synthetic.lorentzian_constants(m["energies"], M, seed=m["seed"]))
        outs.append(torch.empty((len(m["energies"]), 1, ny, nx), dtype=torch.float32, device="cuda"))
        kw = dict(interpolation=m["interpolation"], windowing=m["windowing"], precision=capi.FP32)
        sims.append({p: capi.Sim(m["dims"], M, m["phys_size"], m["energies"], m["angles"], pipeline=p, **kw)
                     for p in (capi.PIPE_PLAIN, capi.PIPE_EBASIS, capi.PIPE_MBASIS)})
        units.append(nx * ny * nz * sims[-1][capi.PIPE_EBASIS].num_angles * len(m["energies"]))
    torch.cuda.synchronize()
    total_units = float(sum(units))
    print(f"{a.count} members, {total_units:.3e} voxel*angle*energy units", file=sys.stderr, flush=True)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    eb = [s[capi.PIPE_EBASIS] for s in sims]
    # shapes tuned once up front (the batch call would do it itself)
    for s in {tuple(x.dims): x for x in eb}.values():
        capi.autotune(s)

    def serial(pipe, idx):
        for i in idx:
            capi.simulate_device(sims[i][pipe], vox[i], mats[i], outs[i])

    res = {}
    # E-basis one at a time (reference outputs for the batch comparison)
    serial(capi.PIPE_EBASIS, range(a.count))
    t = min(timed(lambda: serial(capi.PIPE_EBASIS, range(a.count))) for _ in range(a.reps))
    res["ebasis_serial"] = {"seconds": t, "value": total_units / t}
    ref = [o.cpu().numpy().copy() for o in outs]

    def compare():
        worst, same = 0.0, True
        for o, r in zip(outs, ref):
            g = o.cpu().numpy()
            same &= bool(np.array_equal(np.isnan(g), np.isnan(r)))
            fin = np.isfinite(r)
            for e in range(r.shape[0]):
                f = fin[e]
                den = np.linalg.norm(r[e][f].astype(np.float64))
                if den > 0:
                    worst = max(worst, float(np.linalg.norm(g[e][f].astype(np.float64) - r[e][f]) / den))
        return worst, same

    ps = [s[PIPE] for s in sims]
    if pn != "ebasis":
        serial(PIPE, range(a.count))
        t = min(timed(lambda: serial(PIPE, range(a.count))) for _ in range(a.reps))
        w, same = compare()
        # how much of a one-at-a-time pass the device spends inside kernels: the
        # rest (launch gaps, host work between members) is what removing the gaps
        # between kernels could reclaim; concurrent members can also overlap
        # kernels that do not fill the device alone
        capi.kernel_timing(True)
        t_busy = timed(lambda: serial(PIPE, range(a.count)))
        capi.kernel_timing(False)
        ksum = sum(k["total_us"] for k in capi.kernel_timing_report().values()) * 1e-6
        res[f"{pn}_serial"] = {"seconds": t, "value": total_units / t, "max_rel_l2_vs_ebasis": w,
                               "nan_mask_identical": same, "kernel_seconds": ksum,
                               "kernel_busy_fraction": ksum / t,
                               "note": "kernel_seconds: sum of the launches' HIP-event durations in one serial pass "
                                       f"({t_busy:.4f} s with the events)"}

    capi.simulate_batch(ps, vox, mats, outs)
    t = min(timed(lambda: capi.simulate_batch(ps, vox, mats, outs)) for _ in range(a.reps))
    w, same = compare()
    res[f"{pn}_batch"] = {"seconds": t, "value": total_units / t, "max_rel_l2_vs_ebasis": w,
                          "nan_mask_identical": same}

    t0 = time.perf_counter()
    plan = capi.BatchPlan(ps, vox, mats, outs)
    t_plan = time.perf_counter() - t0
    plan.launch()
    t = min(timed(plan.launch) for _ in range(a.reps))
    w, same = compare()
    res[f"{pn}_graph"] = {"seconds": t, "value": total_units / t, "plan_create_s": t_plan,
                          "max_rel_l2_vs_ebasis": w, "nan_mask_identical": same}
    plan.close()

    # plain one at a time on the first members, scaled by their unit share
    pc = min(a.plain_count, a.count)
    idx = list(range(pc))
    t = timed(lambda: serial(capi.PIPE_PLAIN, idx))
    pu = float(sum(units[i] for i in idx))
    w_plain = 0.0
    for i in idx:
        g, r = outs[i].cpu().numpy(), ref[i]
        fin = np.isfinite(r)
        for e in range(r.shape[0]):
            den = np.linalg.norm(r[e][fin[e]].astype(np.float64))
            if den > 0:
                w_plain = max(w_plain, float(np.linalg.norm(g[e][fin[e]].astype(np.float64) - r[e][fin[e]]) / den))
    res["plain_serial"] = {"seconds_measured": t, "members_measured": pc, "value": pu / t,
                           "max_rel_l2_ebasis_vs_plain": w_plain,
                           "note": "measured on the first members; value = their units / their time"}

    plain_v = res["plain_serial"]["value"]
    out = {
        "metric": "voxel·angle·energy/s (batched vs plain one-at-a-time)", "unit": "voxel·angle·energy/s",
        "workload": f"{a.count} uneven morphologies {{128^2x32, 256^2x64, 384^2x48, 512^2x64}}, 2-4 materials, "
                    "8-32 energies, 18 angles, linear, no window, FP32",
        "total_units": total_units, "modes": res,
        "pipeline": pn,
        "speedup_batch_vs_plain_serial": res[f"{pn}_batch"]["value"] / plain_v,
        "speedup_graph_vs_plain_serial": res[f"{pn}_graph"]["value"] / plain_v,
        "speedup_batch_vs_serial": res[f"{pn}_batch"]["value"] / res[f"{pn}_serial"]["value"],
        "speedup_graph_vs_serial": res[f"{pn}_graph"]["value"] / res[f"{pn}_serial"]["value"],
        "gap_removal_bound_vs_serial": (1.0 / res[f"{pn}_serial"]["kernel_busy_fraction"]
                                        if "kernel_busy_fraction" in res[f"{pn}_serial"] else None),
    }
    del vox, outs
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    sys.exit(main())
