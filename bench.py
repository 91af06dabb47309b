#!/usr/bin/env python3
"""Headline benchmark: Mcell-steps/s of capaz's explicit three-phase FP64
filtration step at 256^3 on one MI355X, with the HBM roofline of the fused
step kernel and the CPU oracle timed beside it.

    python bench.py [--steps K] [--warmup W] [--n 256] [--no-cpu-baseline]

A step is one fused time step (capaz_step: closures, upwind Darcy fluxes,
mass update, 4-iteration 3x3 Newton per cell and the boundary faces) over all
cells of the synthetic config-1 physics (capaz_amd/workloads.py).  The state
is resident in HBM before timing starts.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s (3-phase 256^3 FP64) + %HBM roofline; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Compulsory traffic of one single-pass ping-pong three-phase step per cell:
# read P_w, S_w, S_n, m, K and write P_w, S_w, S_n (8 doubles).
ALGO_BYTES_PER_CELL = 64


def cpu_baseline(nz_slab: int = 64, steps: int = 2) -> dict:
    """The oracle (C restatement of the reference CPU path, one core) on a
    slab of the same workload: same physics and seeded fields, 256 x 256 x
    nz_slab cells, `steps` steps."""
    from capaz_amd import workloads
    from oracle import oracle

    keys, st = workloads.headline(256)
    keys = dict(keys, NZ=nz_slab)
    st = {k: v[:nz_slab].copy() for k, v in st.items()}
    d = oracle.OracleDomain(keys, st)
    t0 = time.perf_counter()
    d.run(steps)
    dt = time.perf_counter() - t0
    cells = 256 * 256 * nz_slab
    return {"value": cells * steps / dt / 1e6, "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of a 256x256x{nz_slab} slab of the headline workload "
                      f"(same physics and seeded fields), oracle/liboracle.so on 1 core, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (a copy of known size, to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        from capaz_amd import hip
        nbytes = args.calibrate_copy << 20
        a, b = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
        hip.lib().hipMemset(a.ptr, 1, nbytes)
        for _ in range(5):
            hip._ok(hip.lib().hipMemcpy(b.ptr, a.ptr, nbytes, hip.D2D), "hipMemcpy")
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 5}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this backend runs one domain on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not a build)", file=sys.stderr)
        return 2

    from capaz_amd import capaz, hip, workloads

    keys, st = workloads.headline(args.n)
    cells = args.n ** 3
    dom = capaz.Domain(keys, st)
    del st
    print(f"bench: {args.n}^3 three-phase, warmup {args.warmup}", file=sys.stderr, flush=True)
    dom.run(args.warmup, first_step=1)
    dom.sync()

    # HIP events on the context's stream, where the fused-step kernels run
    stream = dom.stream()
    ev0, ev1 = hip.Event(), hip.Event()
    dom.sync()
    t0 = time.perf_counter()
    ev0.record(stream)
    dom.run(args.steps, first_step=args.warmup + 1)
    ev1.record(stream)
    dom.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_ms(ev1)
    print(f"bench: {args.steps} steps in {wall:.3f} s (events {gpu_ms:.1f} ms)", file=sys.stderr, flush=True)

    launch_s = gpu_ms / 1e3 / args.steps  # one fused-step kernel per step
    achieved = ALGO_BYTES_PER_CELL * cells / launch_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_fused_step", "algorithmic_bytes_per_cell": ALGO_BYTES_PER_CELL,
            "avg_launch_us": round(launch_s * 1e6, 2)}

    out = {
        "metric": METRIC,
        "value": round(cells * args.steps / wall / 1e6, 1),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"three-phase water/NAPL/gas {args.n}^3 FP64, Corey-2 relperms, lognormal K "
                               f"(ln-sigma 0.5), 4 Newton iterations, top-face injection (config-1 physics)",
                   "cells": cells, "grid": [args.n] * 3, "dt": keys["DT"], "newton_iterations": 4},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline()
        cb["gpu_over_cpu"] = round(out["value"] / cb["value"], 1)
        out["cpu_baseline"] = cb
    dom.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
