#!/usr/bin/env python3
"""Throughput benchmark of the MI355X CyRSoXS pipeline.

Workload (default): realistic bulk morphology, 4 materials on a 512x512x128 grid,
PhysSize 2 nm, Gaussian-filtered softmax volume fractions, filtered S, sphere-
uniform Euler angles; 20 energies (280-299 eV) x 36 E angles (0:5:175), linear
Ewald interpolation, Hanning window, FP32.  A step is one whole simulation
(every energy and angle) with the morphology already resident in HBM.

value = voxel*angle*energy per second of the production pipeline (--pipeline,
default the material-basis executor); the other fused executor (E-basis) and the
plain per-angle pipeline (the reference's structure) are timed beside it and
their patterns compared with the headline's.  The
roofline object reports the dominant kernel's achieved HBM bandwidth: every
launch of the timed steps is bracketed by HIP events on the library stream it
runs on (cyr_kernel_timing), and its algorithmic bytes (inputs read once,
outputs written once) are divided by its average duration.  cpu_baseline times the FP64 CPU oracle on a
bounded sample of the same workload on this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "voxel·angle·energy/s + p50/p99 latency, batched vs plain, at ≤1e-4 rel. L2"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
# workload (material-basis executor), scaled by the counters' reading for a 1 GiB
# negation (fetch x2.0, write x1.0); absent where not measured.
TRAFFIC = {"mbasis": {"xfields": 2.954e9, "fft_y": 3.758e9, "mbasis_zband": 2.017e9, "mbasis_quad": 0.375e9,
                      "rotate_quad": 0.274e9}}
# timing labels of the library (cyr_kernel_timing) -> kernel symbol
KERNEL_OF = {"xfields": "xfields_kernel", "fft_y": "fft_ct_kernel (Y pass)", "fft_z": "fft_ct_kernel (Z pass)",
             "mbasis_extract": "mbasis_extract_kernel", "mbasis_zband": "mbasis_zband_kernel", "mbasis_angles": "mbasis_angles_kernel",
             "mbasis_quad": "mbasis_quad_kernel", "rotate_quad": "rotate_orbit_kernel / rotate_quad_kernel",
             "mbasis_bands": "mbasis_bands_kernel", "ebasis_yz": "ebasis_yz_kernel", "ebasis_dc": "ebasis_dc_kernel",
             "ebasis_angles": "ebasis_angles_kernel", "ebasis_bins": "ebasis_bins_kernel",
             "rotate_all": "rotate_all_kernel"}
PIPE_DESC = {"mbasis": "mbasis (7 energy-independent basis fields per material, packed in pairs: X FFT fused with "
                       "the fields, Y and Z FFT passes, values in the Ewald band; per energy a 2D combination)",
             "ebasis": "ebasis (5 tensor fields per energy, XY FFT, pruned z DFT at the Ewald bins)"}


def _pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    i = max(0, min(len(xs) - 1, int(round(p / 100.0 * len(xs) + 0.5)) - 1))  # nearest rank
    return xs[i]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(vox_t, mats, desc, budget_s):
    """FP64 oracle on 1 energy and as many angles as fit ~budget_s (at least one)."""
    import numpy as np

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O

    log("cpu baseline: copying morphology to host (FP64)")
    vox = vox_t.double().cpu().numpy()
    osim = O.Sim(desc["dims"], desc["num_materials"], desc["phys_size"], desc["energies"][:1], desc["angles"],
                 interpolation=desc["interpolation"], windowing=desc["windowing"])
    nA = osim.num_angles
    nx, ny, nz = desc["dims"]
    V = nx * ny * nz
    done = 0
    t_total = 0.0
    while done < nA:
        t0 = time.perf_counter()
        O.simulate(osim, vox, mats[:1], energies=(0, 1), angles=(done, done + 1))
        t_total += time.perf_counter() - t0
        done += 1
        log(f"cpu baseline: {done} angle(s) in {t_total:.1f} s")
        if t_total > budget_s or t_total / done * (done + 1) > 1.5 * budget_s:
            break
    return {"value": V * done / t_total, "unit": "voxel·angle·energy/s", "cores": O.max_threads(),
            "kind": "port",
            "sample": f"FP64 oracle, 1 energy x {done} of {nA} angles of the same 512x512x128 4-material workload "
                      f"({t_total:.1f} s)"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--pipeline", choices=("mbasis", "ebasis"), default="mbasis")
    ap.add_argument("--other-steps", type=int, default=2,
                    help="timed steps of the other fused executor (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-steps", type=int, default=1,
                    help="steps of the host-buffer entry point (cyr_simulate: morphology uploaded and patterns "
                         "returned over PCIe each call), reported beside the headline; 0 skips it")
    ap.add_argument("--energies", type=int, default=0,
                    help="run only the first N energies (profiling runs; 0 = all 20)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--lanes", type=int, default=1,
                    help="(energy, k) units in flight in the timed steps (the per-kernel roofline pass uses 1)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 1 GiB device-to-device copy (the reference kernel for HBM counter scaling)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported (energies and "
                                   "requests shard across devices without communication; see DESIGN.md)"}))
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import numpy as np
    import torch

    from cyrsoxs_amd import capi, synthetic

    torch.cuda.set_device(0)
    desc = synthetic.config1_desc()
    if a.energies:
        desc["energies"] = desc["energies"][:a.energies]
    dims, M = desc["dims"], desc["num_materials"]
    nx, ny, nz = dims
    V = nx * ny * nz
    log(f"generating morphology {dims} x {M} materials on the GPU")
    vox = synthetic.bulk_morphology(dims, M, seed=1, sigma=6.0, device="cuda")
    mats = synthetic.lorentzian_constants(desc["energies"], M, seed=1)
    torch.cuda.synchronize()
    nE = len(desc["energies"])

    def mk(pipeline):
        return capi.Sim(dims, M, desc["phys_size"], desc["energies"], desc["angles"],
                        interpolation=desc["interpolation"], windowing=desc["windowing"], precision=capi.FP32,
                        pipeline=pipeline)

    if a.copy_calib:
        # a plain elementwise kernel of known traffic: reads 1 GiB, writes 1 GiB
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    pipes = {"mbasis": capi.PIPE_MBASIS, "ebasis": capi.PIPE_EBASIS}
    other = "ebasis" if a.pipeline == "mbasis" else "mbasis"
    sim = mk(pipes[a.pipeline])
    nA = sim.num_angles
    units = V * nA * nE
    out = torch.empty((nE, 1, ny, nx), dtype=torch.float32, device="cuda")

    def run(s, o):
        capi.simulate_device(s, vox, mats, o)

    # one-time per grid shape, like an FFT plan: pick the Y/z launch shape on this device
    t0 = time.perf_counter()
    tuned = capi.autotune(mk(capi.PIPE_EBASIS))
    tune_s = time.perf_counter() - t0
    log(f"autotune (E-basis Y/z): columns {tuned[0]}, target blocks {tuned[1]} ({tune_s:.2f} s)")
    capi.set_energy_lanes(a.lanes)
    log(f"{a.pipeline} pipeline ({a.lanes} lane(s)): warmup {a.warmup}, steps {a.steps}")
    for _ in range(a.warmup):
        run(sim, out)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0 = time.perf_counter()
        run(sim, out)  # returns after the stream has drained
        lat.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = units * a.steps / elapsed
    log(f"{a.pipeline}: {1000.0 * elapsed / a.steps:.2f} ms/step")

    # per-kernel roofline pass: one unit in flight, every launch bracketed by HIP
    # events on the stream it is queued on (the library's stream)
    capi.set_energy_lanes(1)
    capi.kernel_timing(True)
    run(sim, out)
    torch.cuda.synchronize()
    capi.kernel_timing(False)
    kern = capi.kernel_timing_report()
    capi.set_energy_lanes(a.lanes)

    result = {
        "metric": METRIC, "value": value, "unit": "voxel·angle·energy/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1000.0 * elapsed / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian-filtered bulk morphology, Lorentzian optical constants)",
        "config": {"workload": "bulk 512x512x128, 4 materials, PhysSize 2 nm, 20 energies x 36 E angles "
                               "(0:5:175), linear Ewald, Hanning, FP32",
                   "pipeline": PIPE_DESC[a.pipeline],
                   "lanes": a.lanes,
                   "units_per_step": units,
                   "yz_tuning": {"columns": tuned[0], "target_blocks": tuned[1], "autotune_s": round(tune_s, 3)}},
        "latency_ms": {"p50": 1000.0 * _pct(lat, 50), "p99": 1000.0 * _pct(lat, 99)},
    }

    def compare(g, p):
        fin = np.isfinite(p)
        rel = [float(np.linalg.norm((g[e][fin[e]] - p[e][fin[e]])) / np.linalg.norm(p[e][fin[e]]))
               for e in range(nE)]
        return max(rel), bool(np.array_equal(np.isnan(g), np.isnan(p)))

    if a.other_steps > 0:
        osim = mk(pipes[other])
        oout = torch.empty_like(out)
        log(f"{other} pipeline: warmup 1, steps {a.other_steps}")
        run(osim, oout)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.other_steps):
            run(osim, oout)
        torch.cuda.synchronize()
        oel = time.perf_counter() - t0
        rel, same = compare(out.double().cpu().numpy(), oout.double().cpu().numpy())
        result[other] = {"value": units * a.other_steps / oel, "ms_per_step": 1000.0 * oel / a.other_steps,
                         "steps": a.other_steps, "structure": PIPE_DESC[other],
                         f"{a.pipeline}_vs_{other}_max_rel_l2": rel, "nan_mask_identical": same}
        del oout

    if not a.no_plain:
        psim = mk(capi.PIPE_PLAIN)
        pout = torch.empty_like(out)
        log(f"plain pipeline: warmup 1, steps {a.plain_steps}")
        run(psim, pout)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.plain_steps):
            run(psim, pout)
        torch.cuda.synchronize()
        pel = time.perf_counter() - t0
        rel, same = compare(out.double().cpu().numpy(), pout.double().cpu().numpy())
        result["plain"] = {"value": units * a.plain_steps / pel, "ms_per_step": 1000.0 * pel / a.plain_steps,
                           "steps": a.plain_steps,
                           "structure": "per angle: polarization, 3 full 3D FFTs, DC, Ewald, rotate/accumulate",
                           f"speedup_{a.pipeline}_vs_plain": value / (units * a.plain_steps / pel),
                           f"{a.pipeline}_vs_plain_max_rel_l2": rel, "nan_mask_identical": same}

    if a.host_steps > 0:
        # the reference's boundary hands over host buffers: the same simulation from a
        # pageable host morphology ([M][V][4] float32, 2.15 GB) to host patterns
        hv = vox.cpu().numpy()
        log(f"host-buffer entry point: warmup 1, steps {a.host_steps}")
        hout = capi.simulate(sim, hv, mats)
        t0 = time.perf_counter()
        for _ in range(a.host_steps):
            hout = capi.simulate(sim, hv, mats)
        hel = time.perf_counter() - t0
        rel, same = compare(hout.astype(np.float64), out.double().cpu().numpy())
        result["host_buffers"] = {"value": units * a.host_steps / hel, "ms_per_step": 1000.0 * hel / a.host_steps,
                                  "steps": a.host_steps, "h2d_bytes": int(hv.nbytes),
                                  "structure": "cyr_simulate: H2D of the morphology, the same device work, D2H of "
                                               "the patterns (PCIe-inclusive; not the headline value)",
                                  "max_rel_l2_vs_device_api": rel, "nan_mask_identical": same}
        del hv, hout

    # achievable HBM bandwidth on this device: a 1 GiB device copy and a 1 GiB
    # elementwise negation (each reads 1 GiB and writes 1 GiB), best of the two,
    # timed with events on torch's stream (where they run)
    n_copy = 1 << 28
    src = torch.ones(n_copy, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    copy_gbs = 0.0
    for fn in (lambda: dst.copy_(src), lambda: torch.neg(src, out=dst)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        copy_gbs = max(copy_gbs, 2.0 * n_copy * 4 / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9)
    del src, dst

    # ---- roofline of the dominant kernel (largest share of the timed region)
    tot = sum(k["total_us"] for k in kern.values())
    table = {}
    for name, k in sorted(kern.items(), key=lambda kv: -kv[1]["total_us"]):
        gbs = k["avg_bytes"] / (k["avg_us"] * 1e-6) / 1e9 if k["avg_us"] > 0 else 0.0
        table[name] = {"calls": k["calls"], "avg_us": round(k["avg_us"], 2), "share": round(k["total_us"] / tot, 4),
                       "algorithmic_bytes": k["avg_bytes"], "GBs": round(gbs, 1)}
    dom = max(kern, key=lambda n: kern[n]["total_us"])
    ach = table[dom]["GBs"]
    result["roofline"] = {"bound": "hbm", "kernel": KERNEL_OF.get(dom, dom), "achieved": ach, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": TRAFFIC.get(a.pipeline, {}).get(dom),
                          "algorithmic_bytes": kern[dom]["avg_bytes"], "avg_us": kern[dom]["avg_us"],
                          "achievable_copy_GBs": round(copy_gbs, 1), "frac_of_copy": round(ach / copy_gbs, 4),
                          "kernels": table}

    if not a.no_cpu_baseline and rank == 0:
        result["cpu_baseline"] = cpu_baseline(vox, mats, desc, a.cpu_budget)

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
