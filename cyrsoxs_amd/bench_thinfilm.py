"""Thin-film precision study: 3 materials on a 2048 x 2048 x 1 grid, PhysSize 1 nm.

The morphology is a seeded spinodal-like field (2D Gaussian random field with
its spectral peak at q = 2 pi / 40 nm^-1, thresholded into 3 phases with a
2-voxel smooth interface); the director is in-plane and tangent to the phase
boundaries, S = 0.8 at interfaces and 0.2 in the bulk.  50 energies 280.0-304.5
eV in 0.5 eV steps, EAngleRotation 0:2:358 (180 angles), 2D path.

Runs, all with the fused executor chosen by --pipeline (default the
material-basis executor) on the GPU: FP64, FP32 (FP64 twiddle tables, FP64 sums
over angles), FP32 with FP16 morphology storage, FP32 with BF16 storage.  Every FP32
run is compared per pattern against the GPU FP64 run of the same workload (the
GPU FP64 path is held to 1e-10 relative L2 of the FP64 CPU oracle by the test
suite, including a thin-film case).  Per pattern: relative L2 over the detector
mask (pixels finite in the reference) and the largest pixel error relative to
the pattern's largest value.  Also reported: throughput of each run.

    python -m cyrsoxs_amd.bench_thinfilm [--n 2048] [--energies 50] [--pipeline mbasis|ebasis]
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--energies", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pipeline", choices=("mbasis", "ebasis"), default="mbasis")
    a = ap.parse_args(argv)

    import numpy as np
    import torch

    from cyrsoxs_amd import capi, synthetic

    torch.cuda.set_device(0)
    n = a.n
    energies = [280.0 + 0.5 * i for i in range(a.energies)]
    angles = (0.0, 2.0, 358.0)
    vox32 = synthetic.thinfilm_morphology(n, seed=a.seed, device="cuda")
    vox64 = vox32.double()
    mats = synthetic.lorentzian_constants(energies, 3, seed=a.seed)
    units = n * n * 180 * len(energies)

    def run(precision, storage, vox):
        sim = capi.Sim((n, n, 1), 3, 1.0, energies, angles, interpolation=1, precision=precision, storage=storage,
                       pipeline=capi.PIPE_MBASIS if a.pipeline == "mbasis" else capi.PIPE_EBASIS)
        dt = torch.float64 if precision == capi.FP64 else torch.float32
        out = torch.empty((len(energies), 1, n, n), dtype=dt, device="cuda")
        dv = vox
        if storage == capi.STORE_FP16:
            dv = vox.to(torch.float16)
        elif storage == capi.STORE_BF16:
            dv = vox.to(torch.bfloat16)
        capi.simulate_device(sim, dv, mats, out)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        capi.simulate_device(sim, dv, mats, out)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0  # before the patterns are copied to the host
        return out.double().cpu().numpy(), elapsed

    def errors(got, ref):
        rel, mx = [], []
        same = bool(np.array_equal(np.isnan(got), np.isnan(ref)))
        for g, r in zip(got.reshape(-1, n, n), ref.reshape(-1, n, n)):
            f = np.isfinite(r)
            den = np.linalg.norm(r[f])
            rel.append(float(np.linalg.norm(g[f] - r[f]) / den) if den > 0 else 0.0)
            top = np.abs(r[f]).max() if f.any() else 0.0
            mx.append(float(np.abs(g[f] - r[f]).max() / top) if top > 0 else 0.0)
        return {"max_rel_l2": max(rel), "median_rel_l2": float(np.median(rel)), "max_pixel_err_rel_to_peak": max(mx),
                "nan_mask_identical": same}

    ref64, t64 = run(capi.FP64, capi.STORE_NATIVE, vox64)
    res = {"fp64": {"seconds": t64, "value": units / t64}}
    runs = {}
    for name, prec, sto, v in (("fp32", capi.FP32, capi.STORE_NATIVE, vox32), ("fp32_fp16_storage", capi.FP32,
                                                                               capi.STORE_FP16, vox32),
                               ("fp32_bf16_storage", capi.FP32, capi.STORE_BF16, vox32)):
        got, t = run(prec, sto, v)
        runs[name] = got
        res[name] = {"seconds": t, "value": units / t, "vs_gpu_fp64": errors(got, ref64)}
        print(f"{name}: {res[name]}", file=sys.stderr, flush=True)

    out = {"workload": f"thin film {n}x{n}x1, 3 materials, PhysSize 1 nm, {len(energies)} energies 280-"
                       f"{energies[-1]} eV, 180 angles (0:2:358)", "pipeline": a.pipeline, "units": units,
           "results": res}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
