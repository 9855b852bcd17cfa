"""Benchmark of the hydro flux-and-update path on one MI355X.

Workload: 3-D ideal hydro, 256^3 zones in 64^3 blocks (nghost 4), periodic [0,1)^3,
WENO-Z + HLLC, SSP-RK3 (the reference's beta table), CFL 0.8, gamma 5/3, seeded smooth
Fourier-mode state. A step is one full cycle (3 RK stages, ghost fills and the CFL
dt reduction) of the fused FP64 path with all fields resident in HBM.

Prints one JSON line: zone-cycles/s of the FP64 form (the reference's rounding, bitwise
equal to the CPU oracle), the FP32-mixed and FP64-FMA forms beside it with their
relative L1 distance to FP64, a stage-level HBM roofline measured with HIP events on
the simulation's stream, and the CPU oracle timed on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

WORKLOAD = "hydro3d_256c_64b_wenoz_hllc_rk3"
N, NB, NG = 256, 64, 4
SEED = 20261204
HBM_PEAK_GBS = 8000.0
BYTES_PER_ZONE_STAGE = 8 * 3 * 5  # stage input + RK base read, stage output written


def _setup(km, precision="fp64", n=N, nb=NB):
    cfg = km.Config(**{"hydro/reconstruction": "wenoz", "hydro/riemann": "hllc",
                       "parthenon/time/integrator": "rk3", "hydro/cfl": 0.8,
                       "material/single/gamma": 5.0 / 3.0, "kamayan_amd/precision": precision})
    cfg.c.ndim = 3
    lay = km.Layout(3, NG, (nb, nb, nb), (n // nb, n // nb, n // nb))
    desc = km.mesh_desc(lay)
    cell, _ = km.problem_fill("smooth3d", cfg, desc, lay, seed=SEED)
    return cfg, desc, lay, cell


def _run_gpu(km, cell, precision, steps, warmup, profile):
    cfg, desc, lay, _ = _setup(km, precision)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    sim.set_primitive(cell)
    sim.advance(warmup)
    if profile:
        sim.profile(True)
    km.check(km.lib().km_synchronize(), "sync")
    t0 = time.perf_counter()
    sim.advance(steps)
    km.check(km.lib().km_synchronize(), "sync")
    t1 = time.perf_counter()
    times = sim.kernel_times() if profile else None
    state, _ = sim.state()
    sim.close()
    return t1 - t0, times, state, lay


def _rel_l1(a, b, lay):
    sl = (slice(None), slice(0, 5)) + lay.interior()
    x, y = a[sl], b[sl]
    num = np.abs(x - y).sum(axis=(0, 2, 3, 4))
    den = np.abs(y).sum(axis=(0, 2, 3, 4))
    return [float(v) for v in num / den]


def _cpu_baseline(km, cell, cycles):
    from oracle import km_oracle  # the checker, timed as the CPU baseline only

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cfg, desc, lay, _ = _setup(km, "fp64")
    ref = km_oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    t0 = time.perf_counter()
    ref.advance(cycles)
    t1 = time.perf_counter()
    ref.close()
    zones = N ** 3
    return {"value": zones * cycles / (t1 - t0), "unit": "zone-cycles/s", "cores": threads,
            "kind": "port",
            "sample": f"{cycles} cycle(s) of the same {N}^3 workload on the FP64 CPU "
                      f"restatement (OpenMP, {threads} threads), {t1 - t0:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-cycles", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the "
              "sharding design is described in DESIGN.md)", file=sys.stderr)
        return 2
    import kamayan_amd as km
    km.lib()
    if km.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1

    _, _, lay, cell = _setup(km)
    zones = N ** 3
    dt64, times, s64, lay = _run_gpu(km, cell, "fp64", args.steps, args.warmup, profile=True)
    value = zones * args.steps / dt64
    st_ms, st_n = times["stage"]
    stage_ms = st_ms / max(st_n, 1)
    achieved = BYTES_PER_ZONE_STAGE * zones / (stage_ms * 1e-3) / 1e9
    kernels = {k: {"avg_ms": (ms / n if n else 0.0), "launches": n}
               for k, (ms, n) in times.items()}
    out = {
        "metric": "zone-cycles/s on 1 MI355X, FP64 vs FP32-mixed at equal L1 error; % HBM peak",
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt64 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded smooth Fourier-mode state, in-repo generator)",
        "config": {"workload": WORKLOAD, "zones": zones, "block": [NB] * 3, "nghost": NG,
                   "reconstruction": "wenoz", "riemann": "hllc", "integrator": "rk3",
                   "cfl": 0.8, "gamma": 5.0 / 3.0, "seed": SEED, "path": "fused",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one RK stage: sweep_x1 + sweep_x2 + sweep_x3_update + ghost_fill",
                     "algorithmic_bytes_per_launch": BYTES_PER_ZONE_STAGE * zones,
                     "avg_launch_ms": stage_ms},
        "kernels": kernels,
    }
    if not args.no_variants:
        var = {}
        for prec in ("fp32_mixed", "fp64_fma"):
            dtv, _, sv, _ = _run_gpu(km, cell, prec, args.steps, args.warmup, profile=False)
            var[prec] = {"value": zones * args.steps / dtv,
                         "speedup_vs_fp64": dt64 / dtv,
                         "rel_l1_vs_fp64_per_var": _rel_l1(sv, s64, lay)}
        out["variants"] = var
    if not args.no_cpu:
        cpu = _cpu_baseline(km, cell, args.cpu_cycles)
        out["cpu_baseline"] = cpu
        out["gpu_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
