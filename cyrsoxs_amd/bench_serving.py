"""Serving workload: a Poisson stream of simulation requests through the engine.

Each request is a seeded morphology with its grid drawn from {256^2 x 64,
512^2 x 128, 1024^2 x 256}, 2-4 materials (generated as for the bulk workload),
1-50 energies over 280-305 eV, EAngleRotation 0:5:175, linear Ewald, Hanning,
FP32.  Morphologies come from a pool generated once per (grid, materials) and
held in pinned host memory; every request still uploads its morphology, runs,
and downloads its patterns.  Admission is by HBM footprint (cyr_footprint_bytes)
against the engine budget.

For each arrival rate the stream runs for --window seconds; requests still
queued when the drain limit passes are cancelled and counted (their latency is
censored, so p99 under overload is a lower bound).  Latency is submit-to-result
on the host, measured by the engine.  Reported per rate: p50/p99 latency,
completed requests, throughput in voxel*angle*energy/s.  The same stream is
also served FIFO by the plain per-angle executor on one slot at the lowest
rate (--plain-window), and by the served executor (--pipeline, default the
choice between the E-basis and material-basis executors per request) on one
slot (FIFO, no overlap) at one rate, for comparison.

    python -m cyrsoxs_amd.bench_serving [--rates 2,5,10,20,40] [--window 15] [--slots 3] [--pipeline auto|mbasis|ebasis]
                                         [--fifo-rate R (0: none)] [--plain-window S (0: none)]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

GRIDS = ((256, 256, 64), (512, 512, 128), (1024, 1024, 256))


def _pct(xs, p):
    if not xs:
        return None
    s = sorted(xs)
    k = max(0, min(len(s) - 1, int(math.ceil(p / 100.0 * len(s))) - 1))
    return s[k]


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--rates", default="2,5,10,20,40")
    ap.add_argument("--window", type=float, default=15.0)
    ap.add_argument("--drain", type=float, default=20.0)
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--grids", default="256x256x64,512x512x128,1024x1024x256")
    ap.add_argument("--plain-window", type=float, default=10.0)
    ap.add_argument("--fifo-rate", type=float, default=5.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pipeline", choices=("auto", "mbasis", "ebasis"), default="auto")
    return ap.parse_args(argv)


def main(argv=None):
    print(json.dumps(run(parse(argv))), flush=True)
    return 0


def run(a):
    """Serves the Poisson streams; returns the result object (the JSON line)."""
    import numpy as np
    import torch

    from cyrsoxs_amd import capi, synthetic

    torch.cuda.set_device(0)
    PIPE = {"auto": capi.PIPE_AUTO, "mbasis": capi.PIPE_MBASIS, "ebasis": capi.PIPE_EBASIS}[a.pipeline]
    grids = [tuple(int(x) for x in g.split("x")) for g in a.grids.split(",")]
    rates = [float(r) for r in a.rates.split(",")]

    # morphology pool in pinned host memory
    pool = {}
    t0 = time.perf_counter()
    for gi, g in enumerate(grids):
        for M in (2, 3, 4):
            d = synthetic.bulk_morphology(g, M, seed=a.seed * 100 + gi * 10 + M, sigma=6.0, device="cuda")
            h = torch.empty(d.shape, dtype=d.dtype, pin_memory=True)
            h.copy_(d)
            del d
            torch.cuda.empty_cache()
            pool[(g, M)] = h.numpy()
    print(f"pool of {len(pool)} morphologies in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    def make_request(rng, pipeline):
        g = grids[int(rng.integers(0, len(grids)))]
        M = int(rng.integers(2, 5))
        nE = int(rng.integers(1, 51))
        energies = list(np.linspace(280.0, 305.0, nE)) if nE > 1 else [292.5]
        mats = synthetic.lorentzian_constants(energies, M, seed=int(rng.integers(1, 1 << 30)))
        sim = capi.Sim(g, M, 2.0, energies, (0.0, 5.0, 175.0), interpolation=1, windowing=1, precision=capi.FP32,
                       pipeline=pipeline)
        units = g[0] * g[1] * g[2] * sim.num_angles * nE
        return sim, pool[(g, M)], mats, units

    # every grid shape tuned once, as a server would at start-up
    for g in grids:
        capi.autotune(capi.Sim(g, 2, 2.0, [285.0], (0.0, 5.0, 175.0), interpolation=1, windowing=1,
                               pipeline=capi.PIPE_EBASIS))

    def serve(rate, window, slots, pipeline, seed):
        # arrival times and requests drawn in turn, so a shorter window serves a
        # prefix of the same stream
        rng = np.random.default_rng(seed)
        arrivals, reqs = [], []
        t = 0.0
        while True:
            t += rng.exponential(1.0 / rate)
            if t > window:
                break
            arrivals.append(t)
            reqs.append(make_request(rng, pipeline))
        ids, units = [], {}
        with capi.Engine(slots=slots) as eng:
            t_start = time.perf_counter()
            t_note = t_start
            for at, (sim, vox, mats, u) in zip(arrivals, reqs):
                now = time.perf_counter() - t_start
                if at > now:
                    time.sleep(at - now)
                rid = eng.submit(sim, vox, mats)
                ids.append(rid)
                units[rid] = u
                if time.perf_counter() - t_note >= 60.0:  # progress of long windows
                    t_note = time.perf_counter()
                    st = eng.stats()
                    print(f"rate {rate:g}: {t_note - t_start:.0f} s, {st['submitted']} submitted, "
                          f"{st['completed']} completed, {st['queued']} queued", file=sys.stderr, flush=True)
            deadline = time.perf_counter() + a.drain
            for rid in ids:
                left = deadline - time.perf_counter()
                done, _ = eng.poll(rid)
                if not done and left > 0:
                    try:
                        eng.wait(rid, timeout=left)
                    except capi.CyrError:
                        pass
            cancelled = 0
            for rid in ids:
                done, _ = eng.poll(rid)
                if not done:
                    try:
                        eng.cancel(rid)
                        cancelled += 1
                    except capi.CyrError:
                        pass
            for rid in ids:
                eng.wait(rid)
            t_end = time.perf_counter()
            lat, done_units, failed = [], 0.0, 0
            for rid in ids:
                done, st = eng.poll(rid)
                if st == 0:
                    q, r, l = eng.times(rid)
                    lat.append(l)
                    done_units += units[rid]
                elif st != 5:
                    failed += 1
            stats = eng.stats()
        span = t_end - t_start
        return {"rate": rate, "window_s": window, "slots": slots, "submitted": len(ids), "completed": len(lat),
                "cancelled_backlog": cancelled, "failed": failed,
                "p50_ms": None if not lat else 1000 * _pct(lat, 50), "p99_ms": None if not lat else 1000 * _pct(lat, 99),
                "throughput": done_units / span if span > 0 else 0.0, "span_s": span,
                "rejected": stats["rejected"]}

    results = []
    for i, r in enumerate(rates):
        res = serve(r, a.window, a.slots, PIPE, a.seed * 1000 + i)
        print(json.dumps(res), file=sys.stderr, flush=True)
        results.append(res)
    # the same request stream as the continuous-batching run at that rate (or a
    # fresh one if that rate was not run), served FIFO on one slot
    # (a rate or window of 0 leaves the comparison out)
    fifo = plain = None
    if a.fifo_rate > 0:
        fi = rates.index(a.fifo_rate) if a.fifo_rate in rates else 77
        fifo = serve(a.fifo_rate, a.window, 1, PIPE, a.seed * 1000 + fi)
        print(json.dumps(fifo), file=sys.stderr, flush=True)
    if a.plain_window > 0:
        plain = serve(min(rates), a.plain_window, 1, capi.PIPE_PLAIN, a.seed * 1000 + rates.index(min(rates)))
        print(json.dumps(plain), file=sys.stderr, flush=True)
    out = {"metric": "p50/p99 latency and voxel·angle·energy/s under a Poisson request stream",
           "workload": "grids " + a.grids + ", 2-4 materials, 1-50 energies 280-305 eV, EAngle 0:5:175, linear, "
                       "Hanning, FP32",
           "pipeline": a.pipeline, "continuous_batching": results, f"{a.pipeline}_fifo_one_slot": fifo,
           "plain_fifo_one_slot": plain}
    del pool
    return out


if __name__ == "__main__":
    sys.exit(main())
