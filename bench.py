#!/usr/bin/env python3
"""Throughput benchmark of the MI355X CyRSoXS pipeline.

Headline workload: realistic bulk morphology, 4 materials on a 512x512x128 grid,
PhysSize 2 nm, Gaussian-filtered softmax volume fractions, filtered S, sphere-
uniform Euler angles; 20 energies (280-299 eV) x 36 E angles (0:5:175), linear
Ewald interpolation, Hanning window, FP32.  A step is one whole simulation
(every energy and angle) with the morphology resident in HBM when the timed
region starts and the patterns left in HBM (cyr_simulate_device).

value = voxel*angle*energy per second of the production pipeline (--pipeline,
default the material-basis executor).  Beside it, in the same JSON line:
  host_buffers   the same simulation through the host-buffer entry point
                 (cyr_simulate: morphology uploaded over PCIe and patterns
                 returned each call), its rate also at the top level
  ebasis, plain  the other fused executor and the reference's per-angle
                 structure, their patterns compared with the headline's
  batched        the 64-member uneven batch: one at a time (plain and the fused
                 executor), cyr_simulate_batch and the graph-captured plan
  serving        Poisson request streams through the serving engine (two
                 rates), FIFO one-slot and plain FIFO beside them; the top-level
                 latency_ms is the served requests' p50/p99 at the lower rate
  roofline       the dominant kernel's achieved HBM bandwidth: every launch of
                 one simulation bracketed by HIP events on the library stream it
                 runs on (cyr_kernel_timing), algorithmic bytes / average launch
                 time; traffic = its HBM bytes per launch from two child
                 rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by what the
                 same counters read for a 1 GiB negation
  cpu_baseline   the FP64 CPU oracle on this host's cores on a bounded sample
                 (1 energy x the first angles), and the GPU run on the same
                 sample compared with it

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-batch] [--no-serving] [--no-traffic]
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
METRIC = "voxel·angle·energy/s + p50/p99 latency, batched vs plain, at ≤1e-4 rel. L2"
HBM_PEAK_GBS = 8000.0
# timing labels of the library (cyr_kernel_timing) -> kernel symbol
KERNEL_OF = {"xfields": "xfields_kernel", "fft_y": "fft_ct_kernel (Y pass)", "fft_z": "fft_ct_kernel (Z pass)",
             "mbasis_extract": "mbasis_extract_kernel", "mbasis_zband": "mbasis_zband_kernel", "mbasis_angles": "mbasis_angles_kernel",
             "mbasis_quad": "mbasis_quad_kernel", "rotate_quad": "rotate_orbit_kernel / rotate_quad_kernel",
             "mbasis_bands": "mbasis_bands_kernel", "ebasis_yz": "ebasis_yz_kernel", "ebasis_dc": "ebasis_dc_kernel",
             "ebasis_angles": "ebasis_angles_kernel", "ebasis_bins": "ebasis_bins_kernel",
             "rotate_all": "rotate_all_kernel"}
# timing label -> substring of the kernel name in a rocprofv3 counter file (the
# headline's executor launches each of these symbols for that label only)
SYMBOL_OF = {"xfields": "xfields_kernel", "fft_y": "fft_ct_kernel", "mbasis_zband": "mbasis_zband_kernel",
             "mbasis_quad": "mbasis_quad_kernel", "rotate_quad": "rotate_", "ebasis_yz": "ebasis_yz_kernel"}
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


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff}


def cpu_baseline(vox_t, mats, desc, budget_s, gpu_check):
    """FP64 oracle on 1 energy and as many angles as fit ~budget_s (at least one);
    gpu_check(n_angles) runs the GPU on the same sample and returns its patterns."""
    import numpy as np

    sys.path.insert(0, HERE)
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
    avg = None
    while done < nA:
        t0 = time.perf_counter()
        one = O.simulate(osim, vox, mats[:1], energies=(0, 1), angles=(done, done + 1))
        t_total += time.perf_counter() - t0
        avg = one if avg is None else avg + one  # the oracle's mean over the sample: sum / count
        done += 1
        log(f"cpu baseline: {done} angle(s) in {t_total:.1f} s")
        if t_total > budget_s or t_total / done * (done + 1) > 1.5 * budget_s:
            break
    del vox
    ref = avg / done
    got = gpu_check(done).astype(np.float64)
    same = bool(np.array_equal(np.isnan(got), np.isnan(ref)))
    fin = np.isfinite(ref)
    rel = float(np.linalg.norm(got[fin] - ref[fin]) / np.linalg.norm(ref[fin]))
    out = {"value": V * done / t_total, "unit": "voxel·angle·energy/s", "cores": O.max_threads(),
           "kind": "port",
           "sample": f"FP64 oracle, 1 energy x the first {done} of {nA} angles of the same 512x512x128 4-material "
                     f"workload ({t_total:.1f} s), the angle mean formed from single-angle runs",
           "gpu_vs_oracle_rel_l2": rel, "gpu_vs_oracle_nan_mask_identical": same}
    out.update(cpu_info())
    return out


# ---------------------------------------------------------------- HBM traffic

def _counter_rows(d):
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def measure_traffic(label, energies, timeout_s=240):
    """Runs this script's --traffic-probe under rocprofv3 twice (FETCH_SIZE, then
    WRITE_SIZE) and returns the dominant kernel's HBM bytes per launch, each
    counter scaled by its reading for a 1 GiB negation (reads 1 GiB, writes 1 GiB)."""
    sym = SYMBOL_OF.get(label)
    if sym is None:
        return None, f"no counter mapping for kernel label {label}"
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = ["timeout", "-s", "KILL", str(timeout_s), prof, "--pmc", counter, "-d", d, "-o", "run",
                   "--output-format", "csv", "--", sys.executable, os.path.join(HERE, "bench.py"), "--traffic-probe",
                   "--energies", str(energies)]
            log(f"traffic: {counter} pass under rocprofv3")
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            rows = [x for x in _counter_rows(d) if x.get("Counter_Name") == counter]
            kern = [float(x["Counter_Value"]) for x in rows if sym in x.get("Kernel_Name", "")]
            calib = [float(x["Counter_Value"]) for x in rows if "neg" in x.get("Kernel_Name", "").lower()]
            if not kern or not calib:
                tail = " | ".join(r.stdout.strip().splitlines()[-3:]) if r.stdout else ""
                return None, f"{counter}: no samples (rc {r.returncode}; {tail[:300]})"
            # the negation reads and writes 1 GiB; the counters report KiB
            res[counter] = (sum(kern) / len(kern), (1 << 30) / (max(calib) * 1024.0))
    fetch, ff = res["FETCH_SIZE"]
    write, wf = res["WRITE_SIZE"]
    traffic = fetch * 1024.0 * ff + write * 1024.0 * wf
    return {"bytes": traffic, "fetch_kib": fetch, "write_kib": write, "fetch_scale": ff, "write_scale": wf}, None


def traffic_probe(a):
    """Child of measure_traffic: one warm-up and one measured simulation of the
    headline workload, then the 1 GiB calibration negation."""
    import torch

    from cyrsoxs_amd import capi, synthetic

    torch.cuda.set_device(0)
    desc = synthetic.config1_desc()
    if a.energies:
        desc["energies"] = desc["energies"][:a.energies]
    dims, M = desc["dims"], desc["num_materials"]
    vox = synthetic.bulk_morphology(dims, M, seed=1, sigma=6.0, device="cuda")
    mats = synthetic.lorentzian_constants(desc["energies"], M, seed=1)
    sim = capi.Sim(dims, M, desc["phys_size"], desc["energies"], desc["angles"], interpolation=1, windowing=1,
                   precision=capi.FP32, pipeline=capi.PIPE_MBASIS if a.pipeline == "mbasis" else capi.PIPE_EBASIS)
    out = torch.empty((len(desc["energies"]), 1, dims[1], dims[0]), dtype=torch.float32, device="cuda")
    for _ in range(2):
        capi.simulate_device(sim, vox, mats, out)
    torch.cuda.synchronize()
    del vox
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    return 0


# ---------------------------------------------------------------- main

def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--pipeline", choices=("mbasis", "ebasis"), default="mbasis")
    ap.add_argument("--other-steps", type=int, default=2,
                    help="timed steps of the other fused executor (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-steps", type=int, default=3,
                    help="steps of the host-buffer entry point (cyr_simulate: morphology uploaded and patterns "
                         "returned over PCIe each call), reported beside the headline; 0 skips it")
    ap.add_argument("--energies", type=int, default=0,
                    help="run only the first N energies (profiling runs; 0 = all 20)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--lanes", type=int, default=1,
                    help="(energy, k) units in flight in the timed steps (the per-kernel roofline pass uses 1)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 1 GiB elementwise negation (the reference kernel for HBM counter scaling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-member batched leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the Poisson serving leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes (traffic null)")
    ap.add_argument("--serving-window", type=float, default=8.0)
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    if a.traffic_probe:
        return traffic_probe(a)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported (energies and "
                                   "requests shard across devices without communication)"}))
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

    def mk(pipeline, energies=None, angles=None):
        return capi.Sim(dims, M, desc["phys_size"], energies or desc["energies"], angles or desc["angles"],
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
                   "inputs": "device-resident: the morphology is in HBM when the timed region starts and the "
                             "patterns stay in HBM (cyr_simulate_device); the PCIe-inclusive host-buffer rate is "
                             "host_buffers_value",
                   "pipeline": PIPE_DESC[a.pipeline],
                   "lanes": a.lanes,
                   "units_per_step": units,
                   "yz_tuning": {"columns": tuned[0], "target_blocks": tuned[1], "autotune_s": round(tune_s, 3)}},
        "step_ms": {"p50": 1000.0 * _pct(lat, 50), "p99": 1000.0 * _pct(lat, 99)},
    }

    def compare(g, p):
        fin = np.isfinite(p)
        rel = [float(np.linalg.norm((g[e][fin[e]] - p[e][fin[e]])) / np.linalg.norm(p[e][fin[e]]))
               for e in range(g.shape[0])]
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
        hl = []
        for _ in range(a.host_steps):
            t0 = time.perf_counter()
            hout = capi.simulate(sim, hv, mats)
            hl.append(time.perf_counter() - t0)
        hel = sum(hl)
        rel, same = compare(hout.astype(np.float64), out.double().cpu().numpy())
        result["host_buffers_value"] = units * a.host_steps / hel
        result["host_buffers"] = {"value": units * a.host_steps / hel, "ms_per_step": 1000.0 * hel / a.host_steps,
                                  "ms_min": 1000.0 * min(hl), "steps": a.host_steps, "h2d_bytes": int(hv.nbytes),
                                  "structure": "cyr_simulate: H2D of the morphology, the same device work, D2H of "
                                               "the patterns (PCIe-inclusive)",
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
                          "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                          "algorithmic_bytes": kern[dom]["avg_bytes"], "avg_us": kern[dom]["avg_us"],
                          "achievable_copy_GBs": round(copy_gbs, 1), "frac_of_copy": round(ach / copy_gbs, 4),
                          "kernels": table}

    # ---- the 64-member uneven batch: batched vs one at a time
    if not a.no_batch:
        from cyrsoxs_amd import bench_batch
        del out
        torch.cuda.empty_cache()
        log("batched leg: 64 uneven morphologies")
        b = bench_batch.run(bench_batch.parse(["--count", "64", "--reps", "2", "--plain-count", "4"]))
        modes = b["modes"]
        result["batched"] = {
            "workload": b["workload"], "units": b["total_units"],
            "plain_one_at_a_time": modes["plain_serial"]["value"],
            "plain_note": f"plain measured on the first {modes['plain_serial']['members_measured']} members",
            "mbasis_one_at_a_time": modes["mbasis_serial"]["value"],
            "batch": modes["mbasis_batch"]["value"], "graph": modes["mbasis_graph"]["value"],
            "batch_vs_plain": b["speedup_batch_vs_plain_serial"], "graph_vs_plain": b["speedup_graph_vs_plain_serial"],
            "batch_vs_one_at_a_time": b["speedup_batch_vs_serial"], "graph_vs_one_at_a_time": b["speedup_graph_vs_serial"],
            "one_at_a_time_kernel_busy_fraction": modes[f"{b['pipeline']}_serial"].get("kernel_busy_fraction"),
            "gap_removal_bound_vs_one_at_a_time": b["gap_removal_bound_vs_serial"],
            "max_rel_l2_vs_ebasis": max(modes[m]["max_rel_l2_vs_ebasis"] for m in ("mbasis_serial", "mbasis_batch",
                                                                                     "mbasis_graph")),
            "nan_masks_identical": all(modes[m]["nan_mask_identical"] for m in ("mbasis_serial", "mbasis_batch",
                                                                               "mbasis_graph"))}
        out = None

    # ---- Poisson serving: continuous batching vs FIFO
    if not a.no_serving:
        from cyrsoxs_amd import bench_serving
        torch.cuda.empty_cache()
        log("serving leg: Poisson streams at 2 and 10 requests/s")
        w = a.serving_window
        sv = bench_serving.run(bench_serving.parse(
            ["--rates", "2,10", "--window", str(w), "--drain", str(w), "--slots", "3",
             "--grids", "256x256x64,512x512x128", "--plain-window", str(w / 2), "--fifo-rate", "2"]))
        cb = {f"{r['rate']:g}": {k: r[k] for k in ("submitted", "completed", "cancelled_backlog", "p50_ms", "p99_ms",
                                                   "throughput")} for r in sv["continuous_batching"]}
        fifo, plain = sv["auto_fifo_one_slot"], sv["plain_fifo_one_slot"]
        result["serving"] = {
            "workload": "Poisson arrivals, grids {256^2x64, 512^2x128} (the 1024^2x256 grid of the full stream is left "
                        "to cyrsoxs_amd/bench_serving.py), 2-4 materials, 1-50 energies 280-305 eV, EAngle 0:5:175, "
                        f"linear, Hanning, FP32, host morphology uploaded per request, {w:g} s per rate, 3 slots",
            "continuous_batching": cb,
            "fifo_one_slot_2rps": {k: fifo[k] for k in ("completed", "p50_ms", "p99_ms", "throughput")},
            "plain_fifo_one_slot_2rps": {k: plain[k] for k in ("submitted", "completed", "cancelled_backlog", "p50_ms",
                                                               "p99_ms", "throughput")}}
        low = sv["continuous_batching"][0]
        result["latency_ms"] = {"p50": low["p50_ms"], "p99": low["p99_ms"],
                                "source": f"served requests at {low['rate']:g} req/s (arrival to result on the host)"}
    else:
        result["latency_ms"] = {"p50": result["step_ms"]["p50"], "p99": result["step_ms"]["p99"],
                                "source": "timed headline steps"}

    # ---- the dominant kernel's HBM traffic from rocprofv3 counters (child runs)
    if not a.no_traffic and rank == 0:
        t, why = measure_traffic(dom, nE)
        if t:
            result["roofline"]["traffic"] = t["bytes"]
            result["roofline"]["traffic_detail"] = dict(t, ratio_to_algorithmic=t["bytes"] / kern[dom]["avg_bytes"])
        else:
            result["roofline"]["traffic_note"] = why

    if not a.no_cpu_baseline and rank == 0:
        def gpu_check(n_angles):
            inc = desc["angles"][1]
            s = mk(pipes[a.pipeline], energies=desc["energies"][:1], angles=(0.0, inc, inc * (n_angles - 1)))
            o = torch.empty((1, 1, ny, nx), dtype=torch.float32, device="cuda")
            capi.simulate_device(s, vox, mats[:1], o)
            return o.cpu().numpy()

        result["cpu_baseline"] = cpu_baseline(vox, mats, desc, a.cpu_budget, gpu_check)

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
