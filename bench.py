#!/usr/bin/env python3
"""Headline benchmark: Mcell-steps/s of capaz's explicit three-phase FP64
filtration step at 256^3 on one MI355X, with the HBM roofline of the fused
step kernel and the CPU oracle timed beside it.

    python bench.py [--steps K] [--warmup W] [--n 256] [--no-cpu-baseline] [--no-ensemble] [--no-adjoint]
                    [--no-energy] [--no-dropin] [--no-shard]
                    [--only ensemble|adjoint|energy|dropin|shard]

A step is one fused time step (capaz_step: closures, upwind Darcy fluxes,
mass update, 4-iteration 3x3 Newton per cell and the boundary faces) over all
cells of the synthetic three-phase injection workload (capaz_amd/workloads.py).  The state
is resident in HBM before timing starts.  The secondary measurements (batched
ensemble, adjoint, non-isothermal energy step, the drop-in driver, one
z-slab shard) run before the headline's steps.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s (3-phase 256^3 FP64) + %HBM roofline; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Compulsory traffic of one single-pass ping-pong three-phase step per cell:
# read P_w, S_w, S_n, m, K and write P_w, S_w, S_n (8 doubles).
ALGO_BYTES_PER_CELL = 64
# Measured per interior cell of one main-kernel launch at 256^3: HBM bytes
# (rocprofv3 --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors that make
# the counters read a device-to-device copy of known size correctly) and FP64
# flops (64 lanes x (ADD + MUL + TRANS + 2 FMA) of SQ_INSTS_VALU_*_F64).
# profiles/fused_step_3ph_256.md holds the runs and their commands.
TRAFFIC_BYTES_PER_CELL = 74.7
FP64_FLOPS_PER_CELL = 873.1


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, steps: int = 3, nz_slab: int = 32, slab_steps: int = 2) -> dict:
    """The oracle (C restatement of the reference CPU path, bit-identical to
    it in the reference closures) timed on this host: the full n^3 headline
    workload for `steps` steps with its per-cell passes on all OpenMP threads,
    and a 256 x 256 x nz_slab slab on one thread.  Same physics and seeded
    fields as the GPU run; only the step loop is timed."""
    from capaz_amd import workloads
    from oracle import oracle

    def timed(keys, st, nsteps):
        d = oracle.OracleDomain(keys, st)
        t0 = time.perf_counter()
        d.run(nsteps)
        return time.perf_counter() - t0

    keys, st = workloads.headline(n)
    threads = oracle.threads()
    dt_all = timed(keys, st, steps)
    slab_keys = dict(keys, NZ=nz_slab)
    slab = {k: v[:nz_slab].copy() for k, v in st.items()}
    del st
    oracle.set_threads(1)
    try:
        dt_one = timed(slab_keys, slab, slab_steps)
    finally:
        oracle.set_threads(threads)
    one = n * n * nz_slab * slab_steps / dt_one / 1e6
    return {"value": round(n ** 3 * steps / dt_all / 1e6, 2), "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps of the {n}^3 headline workload (same physics and seeded fields), "
                      f"oracle/liboracle.so with its per-cell passes on {threads} OpenMP threads "
                      f"({_cpu_model()}), {dt_all:.1f} s",
            "one_core": {"value": round(one, 3), "cores": 1,
                         "sample": f"{slab_steps} steps of a {n}x{n}x{nz_slab} slab, {dt_one:.1f} s"}}


def bench_ensemble(steps: int = 500, count: int = 64) -> dict:
    """The batched-ensemble workload (capaz_amd.workloads.ensemble: 64 two-phase
    realizations, edges drawn from {32, ..., 96}): `steps` steps of every
    realization as one graph launch of the batch kernels, with the launch
    shape from the autotuner, against the same steps run realization after
    realization with single-domain launches."""
    from capaz_amd import capaz, hip, workloads

    keys, reals = workloads.ensemble(count=count)
    cells = sum(int(np.prod(r["dims"])) for r in reals)
    b = capaz.Batch(keys, reals)
    tune_ms = b.autotune(4)
    b.run(steps)  # warm-up, and the graph of `steps` steps is captured here
    b.sync()
    ev0, ev1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    ev0.record(b.stream())
    b.run(steps)
    ev1.record(b.stream())
    b.sync()
    wall_b = time.perf_counter() - t0
    gpu_b = ev0.elapsed_ms(ev1) / 1e3
    tuning = b.tuning()
    # sequential: one realization at a time, single-domain fused launches
    for d in b.domains:
        d.run(2)
    b.sync()
    t0 = time.perf_counter()
    for d in b.domains:
        d.run(steps)
        d.sync()
    wall_s = time.perf_counter() - t0
    b.close()
    return {"workload": f"{count} two-phase realizations, edges from {{32,48,64,80,96}}, {steps} steps each "
                        f"(lognormal correlated K, ln-sigma 1.0)",
            "cells": cells, "steps": steps,
            "batched": {"value": round(cells * steps / wall_b / 1e6, 1), "unit": "Mcell-steps/s",
                        "ms_per_step": round(wall_b * 1e3 / steps, 4), "gpu_ms_per_step": round(gpu_b * 1e3 / steps, 4),
                        "launch": "one HIP graph of all steps", "tuning": tuning,
                        "autotune_ms_per_step": round(tune_ms, 4)},
            "sequential": {"value": round(cells * steps / wall_s / 1e6, 1), "unit": "Mcell-steps/s",
                           "ms_per_step": round(wall_s * 1e3 / steps, 4),
                           "launch": "capaz_step per realization, one after another"},
            "speedup": round(wall_s / wall_b, 2)}


def bench_adjoint(n: int = 96, nsteps: int = 400, directions: int = 20) -> dict:
    """The permeability/porosity inversion workload (capaz_amd.workloads.
    history_matching: three-phase n^3, nsteps steps, 16 observation columns
    sampled every 50 steps, observations from a seeded true ln K, trial ln K
    at its mean): wall time of the objective alone (capaz_objective) and of
    objective + gradient (capaz_adjoint with its default segment choice, and
    with sqrt(N) checkpoints), and central
    differences of capaz_objective along random directions in (ln K, m),
    each also against the tangent-linear model's exact directional derivative
    (capaz_tangent).  The headline cost is adjoint_over_forward: the gradient
    under sqrt(N) checkpoints (each segment recomputed on a second stream
    beside the backward passes) over the objective alone; the same with one
    stream, and with the whole trajectory recorded by the forward sweep (the
    default when it fits), are reported beside it."""
    from capaz_amd import capaz, workloads

    keys, st, lnK_true, lnK0, ob = workloads.history_matching(n, nsteps)
    t = capaz.Domain(keys, dict(st, K=np.exp(lnK_true)))
    seen_w, seen_n = [], []
    done = 0
    for s in ob["steps"]:
        t.run(int(s) - done, first_step=done + 1)
        done = int(s)
        seen_w.append(t.field("S_w").ravel()[ob["cells"]])
        seen_n.append(t.field("S_n").ravel()[ob["cells"]])
    t.close()
    obs = dict(ob, S_w=np.array(seen_w), S_n=np.array(seen_n))
    m0 = st["m"]
    d = capaz.Domain(keys, dict(st, K=np.exp(lnK0)))

    def reset(lnK, m):
        for f in ("P_w", "S_w", "S_n", "P_n", "P_g"):
            d.set_field(f, st[f])
        d.set_field("K", np.exp(lnK))
        d.set_field("m", m)

    def timed(fn, reps=3):
        best = None
        for _ in range(reps):
            reset(lnK0, m0)
            t0 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return out, best

    J, t_fwd = timed(lambda: d.objective(nsteps, obs))
    # default segment choice: the whole trajectory recorded by the forward
    # sweep when it fits in a quarter of the device memory (no recompute)
    (J_a, gK, gm), t_adj = timed(lambda: d.adjoint(nsteps, obs))
    segment = d.tuning()["adjoint_segment"]
    # sqrt(N) checkpoints, each segment recomputed on the way back (on a
    # second stream while the passes reverse the segment after it; then on
    # the one stream)
    (_, gK1, gm1), t_adj1 = timed(lambda: d.adjoint(nsteps, obs, checkpoint_every=-1))
    d.tune(adjoint_streams=1)
    (_, gK2, gm2), t_adj1s = timed(lambda: d.adjoint(nsteps, obs, checkpoint_every=-1))
    d.tune(adjoint_streams=2)
    d.tune(adjoint_kernel=0)
    _, t_adj_cell = timed(lambda: d.adjoint(nsteps, obs, checkpoint_every=-1))
    d.tune(adjoint_kernel=1)
    _, t_adj_tiled = timed(lambda: d.adjoint(nsteps, obs, checkpoint_every=-1))
    d.tune(adjoint_kernel=2)
    every = int(round(nsteps ** 0.5))
    rng = np.random.default_rng(2024)
    errs = {1e-6: [], 1e-5: [], 1e-4: [], 1e-3: []}
    tan_errs, t_tan = [], None
    for _ in range(directions):
        vK = rng.standard_normal(lnK0.shape)
        vm = 0.02 * rng.standard_normal(lnK0.shape)
        ad = float(np.sum(gK * vK) + np.sum(gm * vm))
        reset(lnK0, m0)
        t0 = time.perf_counter()
        _, dJ = d.tangent(nsteps, obs, vK, vm)
        dt = time.perf_counter() - t0
        t_tan = dt if t_tan is None else min(t_tan, dt)
        tan_errs.append(abs(dJ - ad) / max(abs(dJ), 1e-300))
        for eps in errs:
            reset(lnK0 + eps * vK, m0 + eps * vm)
            jp = d.objective(nsteps, obs)
            reset(lnK0 - eps * vK, m0 - eps * vm)
            jm = d.objective(nsteps, obs)
            fd = (jp - jm) / (2 * eps)
            errs[eps].append(abs(fd - ad) / max(abs(fd), 1e-300))
    d.close()
    two = bench_adjoint_two_phase(n, nsteps)
    return {"workload": f"three-phase {n}^3, {nsteps} steps, 16 observation columns every 50 steps, "
                        f"gradient in per-cell ln K and porosity",
            "two_phase": two,
            "J": J, "J_equal": J == J_a,
            "forward_ms": round(t_fwd * 1e3, 2),
            "adjoint_over_forward": round(t_adj1 / t_fwd, 2), "sqrt_n_checkpoint_every": every,
            "adjoint_sqrt_n_ms": round(t_adj1 * 1e3, 2),
            "adjoint_sqrt_n_one_stream_ms": round(t_adj1s * 1e3, 2),
            "adjoint_sqrt_n_one_stream_over_forward": round(t_adj1s / t_fwd, 2),
            "streams_change_gradient": bool(not (np.array_equal(gK1, gK2) and np.array_equal(gm1, gm2))),
            "adjoint_recorded_ms": round(t_adj * 1e3, 2), "adjoint_recorded_segment": segment,
            "adjoint_recorded_over_forward": round(t_adj / t_fwd, 2),
            "adjoint_sqrt_n_recompute_stencil_per_cell_ms": round(t_adj_cell * 1e3, 2),
            "adjoint_sqrt_n_recompute_stencil_tiled_ms": round(t_adj_tiled * 1e3, 2),
            "checkpointing_changes_gradient": bool(not (np.array_equal(gK, gK1) and np.array_equal(gm, gm1))),
            "tangent": {"ms": round(t_tan * 1e3, 2),
                        "max_rel_err_adjoint_vs_tangent": float(max(tan_errs)),
                        "median_rel_err_adjoint_vs_tangent": float(np.median(tan_errs))},
            "fd": {"directions": directions, "direction": "ln K ~ N(0,1), m ~ 0.02 N(0,1) per cell",
                   **{f"max_rel_err_eps_{eps:g}": float(max(v)) for eps, v in errs.items()},
                   **{f"median_rel_err_eps_{eps:g}": float(np.median(v)) for eps, v in errs.items()}}}


def bench_adjoint_two_phase(n: int = 96, nsteps: int = 400) -> dict:
    """The same inversion with the two-phase Buckley-Leverett physics
    (capaz_amd.workloads.history_matching(phases=2)): the objective alone,
    the gradient under sqrt(N) checkpoints and with the recorded trajectory,
    best of 2 runs each."""
    from capaz_amd import capaz, workloads

    keys, st, lnK_true, lnK0, ob = workloads.history_matching(n, nsteps, phases=2)
    t = capaz.Domain(keys, dict(st, K=np.exp(lnK_true)))
    seen_w, seen_n = [], []
    done = 0
    for s in ob["steps"]:
        t.run(int(s) - done, first_step=done + 1)
        done = int(s)
        Sn = t.field("S_n").ravel()[ob["cells"]]
        seen_w.append(1. - Sn)
        seen_n.append(Sn)
    t.close()
    obs = dict(ob, S_w=np.array(seen_w), S_n=np.array(seen_n))
    d = capaz.Domain(keys, dict(st, K=np.exp(lnK0)))

    def timed(fn, reps=2):
        best = None
        for _ in range(reps):
            for f in ("P_w", "S_n"):
                d.set_field(f, st[f])
            d.set_field("K", np.exp(lnK0))
            t0 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return out, best

    _, t_fwd = timed(lambda: d.objective(nsteps, obs))
    _, t_sq = timed(lambda: d.adjoint(nsteps, obs, checkpoint_every=-1))
    _, t_rec = timed(lambda: d.adjoint(nsteps, obs))
    d.close()
    return {"workload": f"two-phase Buckley-Leverett {n}^3, {nsteps} steps, same observations",
            "forward_ms": round(t_fwd * 1e3, 2), "adjoint_over_forward": round(t_sq / t_fwd, 2),
            "adjoint_sqrt_n_ms": round(t_sq * 1e3, 2), "adjoint_recorded_ms": round(t_rec * 1e3, 2),
            "adjoint_recorded_over_forward": round(t_rec / t_fwd, 2)}


# Compulsory traffic of one single-pass ping-pong step of the non-isothermal
# two-phase workload per cell: read P_w, S_n, T, m and write P_w, S_n, T
# (7 doubles; its K is a scalar of the configuration -- the kernel reads a
# per-cell K array, uniform here, 8 B more).
ENERGY2_ALGO_BYTES_PER_CELL = 56
# Measured per interior cell of one k_energy2_z launch at 256 x 256 x 128, as
# for the three-phase step: HBM bytes (FETCH_SIZE x 2.0 + WRITE_SIZE) and FP64
# flops; profiles/energy2_2ph_256x256x128.md holds the runs and their commands.
ENERGY2_TRAFFIC_BYTES_PER_CELL = 71.7
ENERGY2_FP64_FLOPS_PER_CELL = 1640.8


def bench_energy(nx: int = 256, ny: int = 256, nz: int = 128, steps: int = 20) -> dict:
    """The non-isothermal two-phase workload (capaz_amd.workloads.
    nonisothermal_2ph, 256 x 256 x 128): `steps` fused steps (the
    configuration's energy closure, one kernel per step) with HIP events
    around every launch, for the roofline of that kernel; the same steps with
    the eight entry-point kernels; the oracle on all host threads, 2 steps of
    a 256 x 256 x 32 slab.  Then the reference's own three-phase energy model
    (energy.cpp, its deck and closures) on the same grid, in its six passes
    and as the eight entry points ("three_phase_reference")."""
    from capaz_amd import capaz, hip, workloads
    from oracle import oracle

    cells = nx * ny * nz
    keys, st = workloads.nonisothermal_2ph(nx, ny, nz)
    d = capaz.Domain(keys, st)
    d.tune(energy_kernel=1)
    d.run(3, first_step=1)
    d.sync()
    stream = d.stream()
    ev0, ev1 = hip.Event(), hip.Event()
    d.tune(timing=1)
    d.sync()
    t0 = time.perf_counter()
    ev0.record(stream)
    d.run(steps, first_step=4)
    ev1.record(stream)
    d.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_ms(ev1)
    kern_ms, launches = d.kernel_time()
    d.tune(timing=0)
    d.tune(energy_kernel=0)
    d.run(1, first_step=steps + 4)
    d.sync()
    t0 = time.perf_counter()
    d.run(steps, first_step=steps + 5)
    d.sync()
    wall_seq = time.perf_counter() - t0
    d.close()
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    launch_s = kern_ms / 1e3 / launches
    achieved = ENERGY2_ALGO_BYTES_PER_CELL * interior / launch_s / 1e9
    flops = ENERGY2_FP64_FLOPS_PER_CELL * interior / launch_s / 1e12
    slab, cpu_steps = 32, 2
    o = oracle.OracleDomain(dict(keys, NZ=slab), {k: v[:slab].copy() for k, v in st.items()})
    t1 = time.perf_counter()
    o.run(cpu_steps)
    cpu = time.perf_counter() - t1
    del o
    out = {"workload": f"non-isothermal two-phase {nx}x{ny}x{nz} (Buckley-Leverett physics with the energy "
                       f"equation: c_r 800, rho_r 2000, lambda 1.5, T0 293 K, inflow 353 K at the top, "
                       f"mu(T) = mu0 exp(-0.02 (T - 293))), 4 Newton iterations",
           "path": "one fused kernel per step (energy_kernel 1)",
           "value": round(cells * steps / wall / 1e6, 1), "unit": "Mcell-steps/s",
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": round(ENERGY2_TRAFFIC_BYTES_PER_CELL * interior),
                        "kernel": "k_energy2_z<2>", "algorithmic_bytes_per_cell": ENERGY2_ALGO_BYTES_PER_CELL,
                        "cells_per_launch": interior, "avg_launch_us": round(launch_s * 1e6, 2),
                        "step_us": round(gpu_ms * 1e3 / steps, 2)},
           "fp64_roofline": {"bound": "fp64-valu", "achieved": round(flops, 2), "peak": FP64_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(flops / FP64_PEAK_TFLOPS, 4),
                             "flops_per_cell": ENERGY2_FP64_FLOPS_PER_CELL},
           "entry_point_sequence_ms_per_step": round(wall_seq * 1e3 / steps, 3),
           "cpu": {"value": round(nx * ny * slab * cpu_steps / cpu / 1e6, 2), "unit": "Mcell-steps/s",
                   "cores": oracle.threads(), "kind": "port",
                   "sample": f"{cpu_steps} steps of a {nx}x{ny}x{slab} slab of the same workload, {cpu:.2f} s"}}
    out["gpu_over_cpu"] = round(out["value"] / out["cpu"]["value"], 1)

    # the reference's own three-phase energy model on the same grid
    keys, st = workloads.energy_3ph(nx, ny, nz)
    d = capaz.Domain(keys, st)

    def timed(kernel, first, n=10):
        d.tune(energy_kernel=kernel)
        d.run(2, first_step=first)
        d.sync()
        t0 = time.perf_counter()
        d.run(n, first_step=first + 2)
        d.sync()
        return (time.perf_counter() - t0) / n

    per = timed(1, 1)
    per_seq = timed(0, 13)
    d.close()
    out["three_phase_reference"] = {
        "workload": f"the reference's three-phase + energy model (energy.cpp, its deck and closures) "
                    f"{nx}x{ny}x{nz}, 3 Newton iterations",
        "path": "six passes per step (energy_kernel 1)",
        "value": round(cells / per / 1e6, 1), "unit": "Mcell-steps/s", "ms_per_step": round(per * 1e3, 3),
        "entry_point_sequence_ms_per_step": round(per_seq * 1e3, 3)}
    return out


def bench_dropin(n: int = 256, warmup: int = 5, steps: int = 50) -> dict:
    """The drop-in path as the reference's driver runs it: main.cpp's own
    read_defines, initialization and time_step_function over the adapter
    (capaz_amd/adapter/_bin/capaz_dropin_bench_3ph, built with the drop-in
    drivers), on the reference's Three-phase/defines.ini deck (its van
    Genuchten closures, 7 Newton iterations) at n^3: `steps` time steps timed
    between HIP device synchronisations, once with the adapter running each
    step as one fused capaz_step and once (CAPAZ_DROPIN_FUSED=0) with the
    entry points as called."""
    import subprocess
    import tempfile
    from capaz_amd import workloads

    binary = os.path.join(ROOT, "capaz_amd", "adapter", "_bin", "capaz_dropin_bench_3ph")
    if not os.path.exists(binary):
        return {"skipped": "capaz_dropin_bench_3ph is not built (make -C oracle dropin, which needs the "
                           "reference sources)"}
    deck = {k: v for k, v in workloads.REF_3PH_DECK.items() if k != "MODEL"}
    deck.update(NX=n, NY=n, NZ=n, TIMEX=1.0, SAVE_PLOTS=10 ** 9, PRINT_SCREEN=10 ** 9)
    res = {}
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "defines.ini"), "w") as f:
            f.write("# deck\n#\n" + "".join(f"{k}={v}\n" for k, v in deck.items()))
        for fused in (1, 0):
            env = dict(os.environ, CAPAZ_DROPIN_FUSED=str(fused))
            r = subprocess.run([binary, str(warmup), str(steps)], cwd=wd, env=env, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=600)
            line = [l for l in r.stdout.splitlines() if l.startswith("{")]
            if r.returncode != 0 or not line:
                return {"failed": r.stdout[-2000:]}
            res[fused] = json.loads(line[-1])
    cells = n ** 3
    # the same deck and steps through capaz_step directly (no driver, no adapter)
    from capaz_amd import capaz, deck as deckmod
    keys = dict(deck, MODEL=3)
    with tempfile.TemporaryDirectory() as wd:
        st = deckmod.initial_state(keys, wd, log=lambda *a: None)
    d = capaz.Domain(keys, st)
    d.run(warmup, first_step=1)
    d.sync()
    t0 = time.perf_counter()
    d.run(steps, first_step=warmup + 1)
    d.sync()
    direct = time.perf_counter() - t0
    d.close()
    res["direct"] = {"steps": steps, "seconds": direct}

    def leg(d):
        return {"ms_per_step": round(d["seconds"] * 1e3 / d["steps"], 4),
                "value": round(cells * d["steps"] / d["seconds"] / 1e6, 1), "unit": "Mcell-steps/s"}

    return {"workload": f"the reference driver (main.cpp) over the adapter, Three-phase/defines.ini deck at {n}^3 "
                        f"(van Genuchten closures, 7 Newton iterations), {steps} steps after {warmup}",
            "fused": leg(res[1]), "entry_points": leg(res[0]),
            "speedup": round(res[0]["seconds"] / res[1]["seconds"], 2),
            "capaz_step_same_deck": leg(res["direct"])}


def bench_adjoint_once(n: int = 96, nsteps: int = 50) -> dict:
    """One gradient of the history-matching workload's first `nsteps` steps
    under sqrt(nsteps) checkpoints (capaz_adjoint, segment recompute on the
    second stream), observations at the steps it reaches, after one untimed
    gradient: a short run of every adjoint kernel, for counter collection."""
    from capaz_amd import capaz, workloads

    keys, st, lnK_true, lnK0, ob = workloads.history_matching(n, 400)
    steps = np.array([s for s in ob["steps"] if s <= nsteps])
    obs = dict(ob, steps=steps, S_w=np.full((len(steps), len(ob["cells"])), 0.3),
               S_n=np.full((len(steps), len(ob["cells"])), 0.2))
    d = capaz.Domain(keys, dict(st, K=np.exp(lnK0)))
    d.adjoint(nsteps, obs, checkpoint_every=-1)
    for f in ("P_w", "S_w", "S_n", "P_n", "P_g"):
        d.set_field(f, st[f])
    d.sync()
    t0 = time.perf_counter()
    J, gK, gm = d.adjoint(nsteps, obs, checkpoint_every=-1)
    d.sync()
    sec = time.perf_counter() - t0
    d.close()
    return {"workload": f"history matching at {n}^3, {nsteps} steps, sqrt(N) checkpoints", "J": J,
            "seconds": round(sec, 4), "finite": bool(np.isfinite(gK).all() and np.isfinite(gm).all())}


def bench_shard(n: int = 256, gpus: int = 8, steps: int = 200, warmup: int = 20) -> dict:
    """One GPU's share of the headline workload under a z-slab decomposition
    over `gpus` GPUs, at its real shape: n x n x (n/gpus + 2) cells (the owned
    planes and one ghost plane per neighbour) of the same physics and state,
    `steps` fused steps timed with HIP events on the context's stream.  The
    halo's device-side work per exchange (the two z exchange planes of P_w,
    S_w, S_n, P_n, P_g; the GPU-to-GPU transfer itself needs a second GPU and
    is not run) is what it adds to a step on the GPU's timeline, two ways:
    pack and unpack as separate launches after the step (capaz_load_faces,
    capaz_save_faces); the pack written by the step itself
    (capaz_set_halo_pack) with the unpack (capaz_save_faces) after it; and
    that pack alone, the transfer landing in the ghost planes
    (capaz_ghost_planes: a z plane is contiguous, so there is no unpack)."""
    import ctypes
    from capaz_amd import capaz, hip, workloads

    nz = n // gpus + 2
    keys, st = workloads.headline(n)
    keys = dict(keys, NZ=nz)
    st = {f: np.ascontiguousarray(v[:nz]) for f, v in st.items()}
    d = capaz.Domain(keys, st)
    del st
    d.run(warmup, first_step=1)
    d.sync()
    stream = d.stream()
    ev0, ev1 = hip.Event(), hip.Event()
    ev0.record(stream)
    d.run(steps, first_step=warmup + 1)
    ev1.record(stream)
    d.sync()
    ms = ev0.elapsed_ms(ev1) / steps
    L = capaz.load()
    plane = n * n
    names = ("P_w", "S_w", "S_n", "P_n", "P_g")
    ids = (ctypes.c_int * len(names))(*[capaz.FIELD_ID[f] for f in names])
    buf = hip.DeviceBuffer(8 * plane * 2 * len(names))
    b = ctypes.c_void_p(buf.address)

    def timed(fn, count):
        ev0.record(stream)
        for _ in range(count):
            fn()
        ev1.record(stream)
        d.sync()
        return ev0.elapsed_ms(ev1) * 1e3 / count

    # per step, on the GPU's timeline (each step's launches take longer than
    # the host needs to queue an iteration): the step alone; the step, then
    # the pack and the unpack as separate launches; the step writing the pack
    # itself, then the unpack; the step writing the pack, with no unpack (the
    # transfer lands in the ghost planes, capaz_ghost_planes).  Blocks of the
    # four alternate; the best block of each counts.
    def step_only(t):
        d.run(1, first_step=t)

    def separate(t):
        d.run(1, first_step=t)
        assert L.capaz_load_faces(d._ctx, ids, len(names), 2, b) == 0
        assert L.capaz_save_faces(d._ctx, ids, len(names), 2, b) == 0

    def in_step(t):
        d.run(1, first_step=t)
        assert L.capaz_save_faces(d._ctx, ids, len(names), 2, b) == 0

    first = warmup + steps + 1
    block = max(steps // 5, 1)
    best = {}
    for _ in range(5):
        for name, fn in (("step", step_only), ("separate", separate), ("in_step", in_step),
                         ("in_place", step_only)):
            d.set_halo_pack(buf.address if name in ("in_step", "in_place") else 0)
            t0 = first

            def iteration():
                nonlocal t0
                fn(t0)
                t0 += 1

            best[name] = min(best.get(name, float("inf")), timed(iteration, block))
            first = t0
    d.set_halo_pack(0)
    buf.close()
    d.close()
    owned = n * n * (nz - 2)
    return {"workload": f"one z-slab of the {n}^3 headline over {gpus} GPUs: {n}x{n}x{nz} "
                        f"({nz - 2} owned + 2 ghost planes)",
            "ms_per_step": round(ms, 4), "value": round(owned / (ms / 1e3) / 1e6, 1),
            "unit": "Mcell-steps/s (owned cells)",
            "halo_pack_unpack_us_per_step": round(best["separate"] - best["step"], 2),
            "halo_with_in_step_pack_us_per_step": round(best["in_step"] - best["step"], 2),
            "halo_in_step_pack_no_unpack_us_per_step": round(best["in_place"] - best["step"], 2),
            "halo_bytes_per_step": 2 * len(names) * 8 * plane}


def bench_steady(d, cells: int, warmup: int = 200, steps: int = 1000) -> dict:
    """`warmup` untimed then `steps` timed fused steps of the domain d from its
    first step (HIP events on the context's stream), the main kernel's
    launches timed by the library."""
    from capaz_amd import hip

    d.run(warmup, first_step=1)
    d.sync()
    stream = d.stream()
    ev0, ev1 = hip.Event(), hip.Event()
    d.tune(timing=1)
    d.sync()
    t0 = time.perf_counter()
    ev0.record(stream)
    d.run(steps, first_step=warmup + 1)
    ev1.record(stream)
    d.sync()
    wall = time.perf_counter() - t0
    kern_ms, launches = d.kernel_time()
    d.tune(timing=0)
    return {"warmup": warmup, "steps": steps, "value": round(cells * steps / wall / 1e6, 1),
            "unit": "Mcell-steps/s", "ms_per_step": round(wall * 1e3 / steps, 4),
            "step_us": round(ev0.elapsed_ms(ev1) * 1e3 / steps, 2), "avg_launch_us": round(kern_ms * 1e3 / launches, 2)}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--no-shard", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--only", choices=("ensemble", "adjoint", "adjoint-once", "energy", "shard", "dropin"), default=None,
                    help="run only this secondary measurement and print it")
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
    if args.only:
        # one secondary measurement alone (for profiling), as its own JSON line
        fn = {"ensemble": bench_ensemble, "adjoint": bench_adjoint, "adjoint-once": bench_adjoint_once,
              "energy": bench_energy, "shard": bench_shard, "dropin": bench_dropin}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return 0

    from capaz_amd import capaz, hip, workloads

    # The headline domain is built and uploaded first, then the secondary
    # measurements run, the ensemble's sustained steps last, so that the
    # headline's warm-up and timed steps follow GPU work directly.  After an
    # idle pause the main kernel runs 10-25% slower for its first ~50 steps
    # (its clock settles under load), which a 5-step warm-up would not absorb.
    keys, st = workloads.headline(args.n)
    cells = args.n ** 3
    dom = capaz.Domain(keys, st)
    del st
    extra = {}
    for name, skip, fn in (("shard", args.no_shard, lambda: bench_shard(args.n)),
                           ("dropin", args.no_dropin, bench_dropin), ("adjoint", args.no_adjoint, bench_adjoint),
                           ("energy", args.no_energy, bench_energy), ("ensemble", args.no_ensemble, bench_ensemble)):
        if not skip:
            print(f"bench: {name}", file=sys.stderr, flush=True)
            extra[name] = fn()

    first = 1
    if args.warmup < 100:
        # Short warm-ups: the domain's first 200 + 1000 steps are measured
        # first and reported as "steady"; the warm-up and timed steps below
        # continue the same run with no pause in between (a pause, even a
        # few frees, brings back the first steps' slower clock).
        print("bench: steady", file=sys.stderr, flush=True)
        extra["steady"] = bench_steady(dom, cells)
        first = 1201
    print(f"bench: {args.n}^3 three-phase, warmup {args.warmup}", file=sys.stderr, flush=True)
    dom.run(args.warmup, first_step=first)
    dom.sync()

    # HIP events on the context's stream, where the fused-step kernels run:
    # one pair around the whole timed region and, from the library, one pair
    # around every launch of the step's main kernel.
    stream = dom.stream()
    ev0, ev1 = hip.Event(), hip.Event()
    dom.tune(timing=1)
    dom.sync()
    t0 = time.perf_counter()
    ev0.record(stream)
    dom.run(args.steps, first_step=first + args.warmup)
    ev1.record(stream)
    dom.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_ms(ev1)
    kern_ms, launches = dom.kernel_time()
    dom.tune(timing=0)
    print(f"bench: {args.steps} steps in {wall:.3f} s (events {gpu_ms:.1f} ms, main kernel "
          f"{kern_ms:.1f} ms over {launches} launches)", file=sys.stderr, flush=True)

    # The main kernel updates the interior cells; the boundary faces are a
    # separate small kernel of the same step.
    interior = (args.n - 2) ** 3
    launch_s = kern_ms / 1e3 / launches
    achieved = ALGO_BYTES_PER_CELL * interior / launch_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(TRAFFIC_BYTES_PER_CELL * interior),
            "kernel": "k_fused_z<3,2,64,8,true>", "algorithmic_bytes_per_cell": ALGO_BYTES_PER_CELL,
            "cells_per_launch": interior, "avg_launch_us": round(launch_s * 1e6, 2),
            "step_us": round(gpu_ms * 1e3 / args.steps, 2)}
    fp64 = None
    if FP64_FLOPS_PER_CELL:
        flops = FP64_FLOPS_PER_CELL * interior / launch_s / 1e12
        fp64 = {"bound": "fp64-valu", "achieved": round(flops, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(flops / FP64_PEAK_TFLOPS, 4), "flops_per_cell": FP64_FLOPS_PER_CELL}

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
                               f"(ln-sigma 0.5), 4 Newton iterations, top-face injection of water and gas",
                   "cells": cells, "grid": [args.n] * 3, "dt": keys["DT"], "newton_iterations": 4},
        "roofline": roof,
        "fp64_roofline": fp64,
    }
    out.update(extra)
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(args.n)
        cb["gpu_over_cpu"] = round(out["value"] / cb["value"], 1)
        out["cpu_baseline"] = cb
    dom.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
