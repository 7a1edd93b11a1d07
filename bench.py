"""Benchmark: zone-updates/s of the FP32 MUSCL-Hancock + HLLC hydro step on one MI355X.

Workload: 3D decaying turbulence on a 256^3 periodic grid (gamma 5/3, rho 1, P 1, a seeded
random-phase solenoidal velocity with P(k) ~ k^-2 on 1 <= |k| <= 8 at rms Mach 1), PLM (MC
limiter) + HLLC, CFL 0.4, synthetic initial conditions generated on the host. A "step" is one
full hydro step (clock, predictor, faces, update) over the whole grid, resident in HBM.

Prints ONE JSON line: the fused one-pass path's throughput as `value`, the plain reference-
structured path beside it, the dominant kernel's HBM roofline (64 algorithmic bytes per
zone-update: one read and one write of the 8-float state) timed with HIP events on the sim's
stream, and the reference CPU scheme (the C restatement of RefHydro) timed on a bounded sample
on this host's usable cores (its CPU affinity; the model from /proc/cpuinfo). Every workload
also carries a per-step fraction (64 B per zone-update over the whole step's time) and a CPU
number: RefHydro for the uniform grids (KH 256^2 in full, Sedov 128^3 for 3 steps), the C
restatement of the block step for the AMR blast and RefHydro on each ensemble grid in turn.

The block-AMR blast (root 8^3 blocks of 16^3 cells, up to 3 levels around a seeded-offset sphere
of radius 0.1 holding P = 100 against rho = 1, P = 1, gamma 1.4; 5024 blocks, 4460 leaves, 18.3 M
leaf cells; MC + HLLC, CFL 0.4) is timed beside it: the fused batched path (one persistent kernel
over device leaf lists, graph-captured) against the plain per-level passes, reported under "amr".
The 2D ensemble (64 seeded Riemann quadrant problems) is reported under "ensemble", 2D
Kelvin-Helmholtz at 256^2 (with the CPU RefHydro on the same 100 steps) and 2048^2 under "kh", and the
reference's published uniform grids (KH 256^2 and 2048^2, Sedov 128^3 and 256^3) with its default
scheme under "published_grids", beside its published figures.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--n 256] [--no-cpu] [--amr-steps 20]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import replace

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_UPDATE = 64      # read 8 FP32 conserved values, write 8
# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (scaled by the counters' reading of a 512 MiB device copy: 2.0 and 1.0), for the
# configuration they were measured on: profiles/k_fused_turbulence_256.md
TRAFFIC_MEASURED = {
    ("k_fused_r", 256, (32, 8, 128)): (2 * 501524 + 524304) * 1024,
    ("k_fused_r", 256, (16, 16, 128)): (2 * 505142 + 524306) * 1024,
    ("k_fused_r", 256, (32, 16, 128)): (2 * 544123 + 524296) * 1024,
}


TRAFFIC_SOURCE = "profiles/k_fused_turbulence_256.md"
# HBM bytes per step of the AMR blast's fused path (the ghost shell pass and the march), from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE summed over a step's kernels, scaled as above
AMR_TRAFFIC_MEASURED = {"blast": (int((0.45 + 0.05 + 0.79 + 0.05 + 2.00 + 0.64) * 1e9),
                                   "profiles/k_amr_fused_blast.md: k_amr_shell texels + k_amr_rim + k_amr_fused, "
                                   "HBM read + write per step")}


# VALU instructions (wave64) per launch of the dominant kernel, measured with rocprofv3 --pmc
# SQ_INSTS_VALU for the configuration they were measured on (profiles/k_fused_turbulence_256.md)
VALU_MEASURED = {
    ("k_fused_r", 256, (32, 8, 64)): 286394543,
    ("k_fused_r", 256, (32, 8, 128)): 282960418,
    ("k_fused_r", 256, (16, 16, 64)): 265036818,
    ("k_fused_r", 256, (16, 16, 128)): 261889406,
    ("k_fused_r", 256, (32, 16, 128)): 253433258,
}
# MI355X VALU issue peak for FP32 without packed math: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_LANE_PEAK = 256 * 4 * 16 * 2.4e9


def valu_issue(kname, n, tile, kernel_ms):
    """Issued VALU lane-slots per second of the kernel against the issue peak (a wave64 FP32
    instruction occupies a SIMD for 4 cycles), from the measured instruction count."""
    insts = VALU_MEASURED.get((kname, n, tuple(tile or ())))
    if insts is None or not kernel_ms:
        return None
    rate = insts * 64 / (kernel_ms * 1e-3)
    return {"insts_per_launch": insts, "lane_ops_per_s": rate, "peak": VALU_LANE_PEAK, "frac": rate / VALU_LANE_PEAK}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """The host's usable cores (this process's CPU affinity) and CPU model (lscpu's "Model name",
    from /proc/cpuinfo)."""
    threads = len(os.sched_getaffinity(0))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, model


def step_roofline(cells, ms):
    """64 algorithmic bytes per zone-update over a whole step's time, against the HBM peak."""
    ach = BYTES_PER_UPDATE * cells / (ms * 1e-3)
    return {"achieved": ach / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK, "ms": ms}


def time_steps(sim, steps):
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    return time.perf_counter() - t0


def cpu_baseline(pb, budget_s=10.0):
    """RefHydro (src/reference/hydro.ts:172-458) restated in C, float64, OpenMP over cells, on
    the same 256^3 state: whole-grid updates (primitives, predictor, faces, update; the state is
    not advanced, so every update does the same work) repeated until `budget_s` seconds of CPU
    work have been timed."""
    import numpy as np
    import oracle
    cores, model = host_cpu()
    oracle.RefHydro.set_threads(cores)
    prm = oracle.params(pb.n, ndim=3, dx=pb.dx, gamma=pb.gamma, cfl=pb.cfl, limiter=pb.limiter,
                        isentropic=pb.isentropic, riemann=pb.riemann)
    h = oracle.RefHydro(prm)
    u1, u2 = pb.state(np.float64)
    h.set_cons(u1, u2)
    del u1, u2
    nz = pb.n[2]
    h.step_slab(0, nz)                     # warm: memo, output buffers, page faults
    reps, t = 0, 0.0
    while t < budget_s and reps < 100:
        t0 = time.perf_counter()
        h.step_slab(0, nz)
        t += time.perf_counter() - t0
        reps += 1
    cells = reps * pb.ncell
    # one thread: slabs of 16 planes, in turn, until ~5 s
    oracle.RefHydro.set_threads(1)
    slabs, t1 = 0, 0.0
    while t1 < budget_s / 2 and slabs < nz // 16:
        k0 = 16 * slabs
        t0 = time.perf_counter()
        h.step_slab(k0, k0 + 16)
        t1 += time.perf_counter() - t0
        slabs += 1
    cells1 = slabs * 16 * pb.n[0] * pb.n[1]
    oracle.RefHydro.set_threads(cores)
    return {"value": cells / t, "unit": "zone-updates/s", "cores": cores, "cpu_model": model, "kind": "port",
            "sample": f"{reps} whole-grid updates of RefHydro (C restatement, float64, {cores} OpenMP threads) on "
                      f"the same {pb.n[0]}^3 state ({cells} zone-updates, {t:.1f} s)",
            "single_thread": {"value": cells1 / t1, "cores": 1,
                              "sample": f"{slabs} slabs of 16 planes, one thread ({cells1} zone-updates, {t1:.1f} s)"}}


def cpu_baseline_steps(pb, steps=100):
    """RefHydro on a uniform grid, run for `steps` steps with dt from the state's own rates
    (src/reference/hydro.ts:172-458 restated in C, float64) at all usable cores, and its first
    fifth at one thread."""
    import numpy as np
    import oracle
    cores, model = host_cpu()
    out = {}
    for threads, nsteps in ((cores, steps), (1, max(steps // 5, 1))):
        oracle.RefHydro.set_threads(threads)
        prm = oracle.params(pb.n, ndim=pb.ndim, dx=pb.dx, gamma=pb.gamma, cfl=pb.cfl, limiter=pb.limiter,
                            isentropic=pb.isentropic, riemann=pb.riemann)
        h = oracle.RefHydro(prm)
        h.set_cons(*pb.state(np.float64))
        t0 = time.perf_counter()
        for _ in range(nsteps):
            h.step()
        t = time.perf_counter() - t0
        grid = "x".join(str(v) for v in pb.n[:pb.ndim])
        out[threads] = {"value": pb.ncell * nsteps / t, "unit": "zone-updates/s", "cores": threads, "kind": "port",
                        "cpu_model": model,
                        "sample": f"{nsteps} steps of RefHydro (C restatement, float64) on the {grid} grid, "
                                  f"{threads} thread(s), {t:.2f} s"}
    oracle.RefHydro.set_threads(cores)
    res = out[cores]
    res["single_thread"] = out[1]
    return res


def cpu_baseline_amr(blocks, root, dx0, u1, u2, cfg, ndim=3, name="blast", steps=1):
    """The block step on the CPU: `steps` steps of the whole tree through the C restatement of
    the block update the GPU runs (predict, faces, coarse-fine means and update of
    src/amr/predict.wgsl and update.wgsl; FP32, OpenMP over blocks). The reference's own CPU
    path (RefHydro) has no blocks, so this restatement stands in for it here."""
    import numpy as np
    import oracle
    cores, model = host_cpu()
    oracle.RefHydro.set_threads(cores)
    t = oracle.AmrTable(blocks, root, ndim=ndim)
    prm = oracle.params(cfg.n, ndim=ndim, dx=dx0, gamma=cfg.gamma, cfl=cfg.cfl, limiter=cfg.limiter,
                        isentropic=cfg.isentropic)
    a1, a2 = u1.copy(), u2.copy()
    clk = np.zeros(4, np.float32)
    t0 = time.perf_counter()
    for _ in range(steps):
        oracle.wgsl32.amr_step(prm, t, dx0, a1, a2, clk)
    dt = time.perf_counter() - t0
    cells = int(oracle_leaf_cells(blocks, ndim)) * steps
    return {"value": cells / dt, "unit": "zone-updates/s", "cores": cores, "cpu_model": model, "kind": "port",
            "sample": f"{steps} step(s) of the whole {name} ({cells} leaf zone-updates) through the C restatement of "
                      f"the block step (FP32, {cores} OpenMP threads over blocks), {dt:.2f} s"}


def oracle_leaf_cells(blocks, ndim=3):
    from shaderous_amd import amr_tree
    return amr_tree.leaf_mask(blocks).sum() * 16 ** ndim


def cpu_baseline_ensemble(ns, states, cfg):
    """RefHydro (C restatement, float64) one step on each of the ensemble's grids in turn at all
    usable cores: the reference runs one simulation at a time."""
    import numpy as np
    import oracle
    cores, model = host_cpu()
    oracle.RefHydro.set_threads(cores)
    t, cells = 0.0, 0
    for n, (u1, u2) in zip(ns, states):
        prm = oracle.params((n, n, 1), ndim=2, dx=1.0 / n, gamma=cfg.gamma, cfl=cfg.cfl, limiter=cfg.limiter,
                            isentropic=cfg.isentropic, bc_lo=cfg.bc_lo, bc_hi=cfg.bc_hi)
        h = oracle.RefHydro(prm)
        h.set_cons(u1.astype(np.float64), u2.astype(np.float64))
        t0 = time.perf_counter()
        h.step()
        t += time.perf_counter() - t0
        cells += n * n
    return {"value": cells / t, "unit": "zone-updates/s", "cores": cores, "cpu_model": model, "kind": "port",
            "sample": f"one step of RefHydro on each of the {len(ns)} grids in turn ({cells} zone-updates, "
                      f"{cores} OpenMP threads), {t:.2f} s"}


def amr_bench(args):
    """The AMR blast: fused batched path vs the plain per-level passes (and the batched passes if
    asked), leaf zone-updates/s and the fused kernel's HBM roofline fraction (64 B per leaf
    zone-update)."""
    from shaderous_amd import amr_tree, problems
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    t0 = time.perf_counter()
    blocks, dx0, prims, centre = problems.amr_blast(seed=7)
    u1, u2 = amr_tree.blocks_state(blocks, dx0, prims, 1.4)
    leaves = int(amr_tree.leaf_mask(blocks).sum())
    cells = leaves * 4096
    log(f"[bench] AMR blast: {len(blocks)} blocks, {leaves} leaves ({cells} cells), IC in {time.perf_counter() - t0:.1f} s")
    cfg = SimConfig(n=(128, 128, 128), ndim=3, dx=dx0, gamma=1.4, limiter="mc", isentropic=False, cfl=0.4)
    res = {}
    for path in args.amr_paths.split(","):
        sim = AmrSim(cfg, blocks, (8, 8, 8), path=path)
        try:
            sim.set_state(u1, u2)
            sim.step(3)
            sim.sync()
            t0 = time.perf_counter()
            sim.step(args.amr_steps)
            sim.sync()
            t = time.perf_counter() - t0
            prof = sim.profile(3)
            st = sim.stats()
            d = sim.diagnostics()
        finally:
            sim.destroy()
        res[path] = {"value": cells * args.amr_steps / t, "ms_per_step": t / args.amr_steps * 1e3, "passes_ms": prof,
                     "launches_per_step": st["launches_per_step"], "mass": d.mass, "e_tot": d.e_tot}
        log(f"[bench] AMR {path}: {t / args.amr_steps * 1e3:.3f} ms/step, {cells * args.amr_steps / t:.3e} upd/s, {prof}")
    out = {"workload": "3D blast on 16^3 blocks: root 8^3, 3 levels, sphere r 0.1 at a seeded offset, P 100 vs 1, "
                       "rho 1, gamma 1.4, PLM (MC) + HLLC, CFL 0.4, static mesh",
           "blocks": len(blocks), "leaves": leaves, "leaf_cells": cells, "steps": args.amr_steps, "paths": res}
    if args.time_bins > 0:
        # per-block timesteps on the passes path: simulated time per wall second against single
        # steps of the same path, and the block updates saved
        rates = {}
        for k in (0, args.time_bins):
            sim = AmrSim(cfg, blocks, (8, 8, 8), path="passes", time_bins=k)
            try:
                sim.set_state(u1, u2)
                sim.step(1 << args.time_bins)
                sim.sync()
                c0 = sim.clock()
                t0 = time.perf_counter()
                sim.step(2 << args.time_bins)
                sim.sync()
                t = time.perf_counter() - t0
                c1 = sim.clock()
                _, work = sim.bins()
            finally:
                sim.destroy()
            rates[k] = {"sim_time_per_s": ((c1[1] + c1[0]) - (c0[1] + c0[0])) / t, "substeps": 2 << args.time_bins,
                        "ms_per_substep": t / (2 << args.time_bins) * 1e3, "work": work}
        kb = rates[args.time_bins]
        out["time_bins"] = {"K": args.time_bins, "single": rates[0], "bins": kb,
                            "update_saving": kb["work"]["single"] / kb["work"]["updates"],
                            "speedup": kb["sim_time_per_s"] / rates[0]["sim_time_per_s"]}
        log(f"[bench] AMR time bins K={args.time_bins}: {out['time_bins']['update_saving']:.2f}x fewer block updates, "
            f"{out['time_bins']['speedup']:.2f}x the simulated time per second of single steps (passes path)")
    if "fused" in res:
        f = res["fused"]
        step_ms = f["passes_ms"]["step"]
        kernel_ms = f["passes_ms"]["update"]
        ach = BYTES_PER_UPDATE * cells / (step_ms * 1e-3)
        out["value"] = f["value"]
        # per step: the ghost shell pass and the march (every launch of a step), by HIP events
        out["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": ach / HBM_PEAK, "per": "step (shell + march)", "step_ms": step_ms,
                           "traffic": AMR_TRAFFIC_MEASURED["blast"][0], "traffic_source": AMR_TRAFFIC_MEASURED["blast"][1],
                           "march": {"kernel": "k_amr_fused", "kernel_ms": kernel_ms,
                                     "frac": BYTES_PER_UPDATE * cells / (kernel_ms * 1e-3) / HBM_PEAK}}
        if "levels" in res:
            out["fused_over_plain_levels"] = f["value"] / res["levels"]["value"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_amr(blocks, (8, 8, 8), dx0, u1, u2, cfg)
        log(f"[bench] AMR CPU: {out['cpu_baseline']['value']:.3e} upd/s")
    return out


def ensemble_bench(args):
    """64 seeded 2D four-quadrant Riemann problems, sizes drawn from {64, 128, 256, 512}^2, gamma
    1.4, outflow, MC + HLLC, CFL 0.4, each with its own dt, advanced together for `ens_steps`
    steps by one persistent launch per step (autotuned tile)."""
    from shaderous_amd import problems
    from shaderous_amd.ensemble import Ensemble
    from shaderous_amd.sim import SimConfig
    ns = problems.ensemble_sizes(64, seed=42)
    cells = sum(n * n for n in ns)
    cfg = SimConfig(n=(64, 64, 1), ndim=2, dx=1.0 / 64, gamma=1.4, bc_lo=(1, 1, 0), bc_hi=(1, 1, 0), limiter="mc",
                    isentropic=False, cfl=0.4)
    if args.tile2d:
        cfg = replace(cfg, tile=(*args.tile2d, 1))
    ens = Ensemble(cfg, [(n, n) for n in ns])
    states = [problems.riemann_quadrants(n, seed=1000 + i).state() for i, n in enumerate(ns)]
    try:
        for i, st in enumerate(states):
            ens.set_state(i, *st)
        tile = tuple(args.tile2d) if args.tile2d else ens.autotune(reps=3)
        ens.step(5)
        ens.sync()
        t0 = time.perf_counter()
        ens.step(args.ens_steps)
        ens.sync()
        t = time.perf_counter() - t0
        prof = ens.profile(5)
    finally:
        ens.destroy()
    value = cells * args.ens_steps / t
    log(f"[bench] ensemble: 64 instances, {cells} cells, tile {tile}: {t / args.ens_steps * 1e3:.3f} ms/step, "
        f"{value:.3e} upd/s, {prof}")
    ach = BYTES_PER_UPDATE * cells / (prof["fused"] * 1e-3)
    # 60-cell-wide tiles are the row march's strips, the others the LDS tile kernel's
    kernel = "k_ens_rows" if tile[0] == 60 else "k_ens_fused"
    out = {"workload": "64 2D four-quadrant Riemann problems, sizes from {64,128,256,512}^2 (seed 42), gamma 1.4, "
                       "outflow, PLM (MC) + HLLC, CFL 0.4, per-instance dt",
           "cells": cells, "steps": args.ens_steps, "value": value, "ms_per_step": t / args.ens_steps * 1e3,
           "tile": tile, "passes_ms": prof,
           "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": ach / HBM_PEAK, "kernel": kernel, "kernel_ms": prof["fused"]},
           "per_step": step_roofline(cells, t / args.ens_steps * 1e3)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ensemble(ns, states, cfg)
        log(f"[bench] ensemble CPU: {out['cpu_baseline']['value']:.3e} upd/s")
    return out


# The reference's published uniform-grid throughputs (zone-updates/s; FP32, its default Koren
# limiter and HLLC, one MUSCL-Hancock stage per step; its own hardware):
# paper/paper/figures/data/performance/throughput_h200.json:1, throughput_h100.json:1,
# throughput_a100.json:1 and throughput.json:1 (RTX 3080)
PUBLISHED = {
    ("kh", 256): {"H200 NVL": 7.25e8, "H100 SXM": 6.22e8, "A100 SXM4": 4.49e8, "RTX 3080": 5.04e8},
    ("kh", 2048): {"H200 NVL": 2.74e9, "H100 SXM": 2.60e9, "A100 SXM4": 1.74e9, "RTX 3080": 1.85e9},
    ("sedov", 128): {"H200 NVL": 1.42e9, "H100 SXM": 1.34e9, "A100 SXM4": 1.04e9, "RTX 3080": 1.26e9},
    ("sedov", 256): {"H200 NVL": 1.67e9, "H100 SXM": 1.57e9, "A100 SXM4": 1.20e9, "RTX 3080": 1.39e9},
    ("kh_amr", 4096): {"H200 NVL": 1.34e9, "H100 SXM": 1.02e9, "A100 SXM4": 8.14e8, "RTX 3080": 9.86e8},
}


def uniform_run(args, pb, steps):
    """Zone-updates/s of a uniform problem on the fused path with the autotuned tile (or
    --tile2d on a flat grid), and the dominant kernel's time by HIP events."""
    from shaderous_amd.sim import Sim, config_for
    flat = pb.n[2] == 1
    forced = (*args.tile2d, 1) if (args.tile2d and flat) else (0, 0, 0)
    with Sim(config_for(pb, path="fused", tile=forced)) as sim:
        sim.set_state(*pb.state())
        tile = forced if forced[0] else sim.autotune(reps=5)
        sim.step(args.warmup)
        t = time_steps(sim, steps)
        prof = sim.profile(5)
        kname = sim.kernel_name()
    value = pb.ncell * steps / t
    return {"value": value, "ms_per_step": t / steps * 1e3, "steps": steps, "tile": list(tile[:2] if flat else tile),
            "kernel": kname, "kernel_ms": prof["update"],
            "hbm_frac": BYTES_PER_UPDATE * pb.ncell / (prof["update"] * 1e-3) / HBM_PEAK,
            "per_step": step_roofline(pb.ncell, t / steps * 1e3)}


def kh_bench(args):
    """2D Kelvin-Helmholtz at 256^2 (the reference-CPU-runnable case, with the CPU RefHydro
    running the same 100 steps) and 2048^2: unit square, periodic, gamma 5/3, PLM (MC) + HLLC,
    CFL 0.4, on the fused path (a flat FP32 grid steps through the 2D kernels)."""
    from shaderous_amd import problems
    out = {}
    for n in (256, 2048):
        pb = problems.kelvin_helmholtz(n)
        r = uniform_run(args, pb, args.kh_steps if n == 256 else max(args.kh_steps // 2, 1))
        log(f"[bench] KH {n}^2: tile {r['tile']} ({r['kernel']}), {r['ms_per_step']:.4f} ms/step, {r['value']:.3e} upd/s")
        if n == 256 and not args.no_cpu:
            r["cpu_baseline"] = cpu_baseline_steps(pb, 100)
            log(f"[bench] KH 256^2 CPU: {r['cpu_baseline']['value']:.3e} upd/s")
        out[f"{n}^2"] = r
    return {"workload": "2D Kelvin-Helmholtz, unit square, periodic, gamma 5/3, rho 2 / 1, vx +-0.5, P 2.5, "
                        "seeded vy noise, PLM (MC) + HLLC, CFL 0.4", "grids": out}


def published_bench(args):
    """The reference's published uniform grids with its default scheme (Koren + isentropic
    reconstruction, HLLC; CFL 0.4): 2D Kelvin-Helmholtz at 256^2 and 2048^2 (the shear layer
    above) and the 3D Sedov blast at 128^3 and 256^3 (gamma 1.4, P 1e-5 around E = 1 within 3
    cells, periodic), on the fused path with the autotuned tile, beside its published figures
    for the same grids on its own hardware."""
    from shaderous_amd import problems
    out = {}
    for kind, n in (("kh", 256), ("kh", 2048), ("sedov", 128), ("sedov", 256)):
        if kind == "kh":
            pb = problems.kelvin_helmholtz(n, limiter="koren", isentropic=True)
            steps = args.kh_steps if n == 256 else max(args.kh_steps // 2, 1)
            name = f"KH {n}^2"
        else:
            pb = problems.sedov(n, limiter="koren", isentropic=True, t_end=math.inf)
            steps = max(args.kh_steps // 2, 1) if n == 128 else max(args.kh_steps // 5, 1)
            name = f"Sedov {n}^3"
        r = uniform_run(args, pb, steps)
        r["reference_published"] = PUBLISHED[(kind, n)]
        r["over_h200"] = r["value"] / PUBLISHED[(kind, n)]["H200 NVL"]
        if kind == "sedov" and n == 128 and not args.no_cpu:
            r["cpu_baseline"] = cpu_baseline_steps(pb, 3)
            log(f"[bench] Sedov 128^3 CPU: {r['cpu_baseline']['value']:.3e} upd/s")
        log(f"[bench] {name} (Koren): tile {r['tile']} ({r['kernel']}), {r['ms_per_step']:.4f} ms/step, "
            f"{r['value']:.3e} upd/s, {r['over_h200']:.1f}x the reference on an H200")
        out[name] = r
    return {"scheme": "PLM (Koren, isentropic reconstruction) + HLLC, CFL 0.4, FP32", "grids": out}


def kh_amr_bench(args):
    """The reference's adaptive KH figure's shape on 2D blocks: the shear layer on a quadtree of
    256^2 root cells refined 4 times (finest 4096^2; problems.kh_adaptive), static mesh, with the
    reference's default scheme (Koren + isentropic reconstruction, HLLC, CFL 0.4, FP32), on the
    batched passes (each pass one launch over every level) and the per-level passes; leaf
    zone-updates/s beside the reference's published figure for its 4096-finest, 4-level run."""
    from shaderous_amd import amr_tree, problems
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    blocks, dx0, prims = problems.kh_adaptive()
    u1, u2 = amr_tree.blocks_state(blocks, dx0, prims, 5.0 / 3.0, ndim=2)
    leaves = int(amr_tree.leaf_mask(blocks).sum())
    cells = leaves * 256
    log(f"[bench] KH adaptive: {len(blocks)} blocks, {leaves} leaves ({cells} cells)")
    cfg = SimConfig(n=(256, 256, 1), ndim=2, dx=dx0, gamma=5.0 / 3.0, limiter="koren", isentropic=True, cfl=0.4)
    res = {}
    for path in ("passes", "levels"):
        sim = AmrSim(cfg, blocks, (16, 16, 1), path=path)
        try:
            sim.set_state(u1, u2)
            sim.step(5)
            sim.sync()
            t0 = time.perf_counter()
            sim.step(args.kh_amr_steps)
            sim.sync()
            t = time.perf_counter() - t0
            prof = sim.profile(5)
            st = sim.stats()
        finally:
            sim.destroy()
        ms = t / args.kh_amr_steps * 1e3
        res[path] = {"value": cells * args.kh_amr_steps / t, "ms_per_step": ms, "passes_ms": prof,
                     "launches_per_step": st["launches_per_step"],
                     "per_step": step_roofline(cells, prof["step"])}
        log(f"[bench] KH adaptive {path}: {ms:.4f} ms/step, {cells * args.kh_amr_steps / t:.3e} upd/s, {prof}")
    pub = PUBLISHED[("kh_amr", 4096)]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_amr(blocks, (16, 16, 1), dx0, u1, u2, cfg, ndim=2, name="adaptive KH", steps=3)
        log(f"[bench] KH adaptive CPU: {cpu['value']:.3e} upd/s")
    return {"workload": "2D Kelvin-Helmholtz shear layer on 16^2 blocks: root 256^2, 4 levels (finest 4096^2), "
                        "static mesh around the interfaces, gamma 5/3, PLM (Koren, isentropic reconstruction) + HLLC, "
                        "CFL 0.4, FP32",
            "blocks": len(blocks), "leaves": leaves, "leaf_cells": cells, "steps": args.kh_amr_steps,
            "value": res["passes"]["value"], "paths": res, "reference_published": pub,
            "reference_leaf_cells": 772864, "over_h200": res["passes"]["value"] / pub["H200 NVL"],
            "batched_over_levels": res["passes"]["value"] / res["levels"]["value"], "cpu_baseline": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tile", default=None, metavar="TX,TY,ZC",
                    help="fused tile and z chunk to use instead of autotuning (e.g. 32,8,64)")
    ap.add_argument("--paths", default="fused,plain", help="which uniform-grid paths to time")
    ap.add_argument("--amr-steps", type=int, default=20, help="timed steps of the AMR blast (0: skip it)")
    ap.add_argument("--amr-paths", default="fused,levels", help="AMR paths to time (fused, passes, levels)")
    ap.add_argument("--time-bins", type=int, default=3, help="per-block timesteps on the AMR blast (0: skip)")
    ap.add_argument("--ens-steps", type=int, default=200, help="steps of the 2D ensemble (0: skip it)")
    ap.add_argument("--kh-amr-steps", type=int, default=100, help="steps of the adaptive 2D KH on blocks (0: skip it)")
    ap.add_argument("--kh-steps", type=int, default=100,
                    help="steps of 2D KH at 256^2 (half at 2048^2 and Sedov 128^3, a fifth at 256^3; 0: skip)")
    ap.add_argument("--tile2d", default=None, metavar="TX,TY", type=lambda v: tuple(int(x) for x in v.split(",")),
                    help="2D tile for the ensemble and KH instead of autotuning (e.g. 60,16)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a device copy of MIB MiB three times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the hydro path is built and measured "
                                   "on a single MI355X (multi-GPU sharding is described, not run)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from shaderous_amd import problems
    from shaderous_amd.sim import Sim, config_for

    n = args.n
    t0 = time.perf_counter()
    pb = problems.turbulence(n, limiter="mc", isentropic=False, cfl=0.4)
    u1, u2 = pb.state()
    log(f"[bench] initial conditions {n}^3 in {time.perf_counter() - t0:.1f} s")
    ncell = pb.ncell

    paths = args.paths.split(",")
    if args.copy_probe:
        import torch
        nb = args.copy_probe << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x {args.copy_probe} MiB device copies")

    value = plain_value = None
    kernel_ms = achieved = kname = None
    tile = None
    diag = None
    t = tp = None
    pprof = None
    # fused one-pass path
    if "fused" in paths:
        forced = tuple(int(v) for v in args.tile.split(",")) if args.tile else None
        sim = Sim(config_for(pb, path="fused", tile=forced or (0, 0, 0)))
        sim.set_state(u1, u2)
        if forced:
            tile = forced
            log(f"[bench] fused tile {tile}")
        elif not args.no_tune:
            tile = sim.autotune(reps=5)
            log(f"[bench] autotuned fused tile {tile}")
        sim.step(args.warmup)
        sim.sync()
        log("[bench] fused warmup done")
        t = time_steps(sim, args.steps)
        value = ncell * args.steps / t
        prof = sim.profile(5)
        kernel_ms = prof["update"]
        kname = sim.kernel_name()
        achieved = BYTES_PER_UPDATE * ncell / (kernel_ms * 1e-3)
        log(f"[bench] fused: {t / args.steps * 1e3:.3f} ms/step, kernel {kernel_ms:.3f} ms, {value:.3e} upd/s")
        diag = sim.diagnostics()
        sim.destroy()

    # plain reference-structured path (rates, clock, predict, tiled update)
    if "plain" in paths:
        sim = Sim(config_for(pb, path="plain"))
        sim.set_state(u1, u2)
        sim.step(args.warmup)
        sim.sync()
        tp = time_steps(sim, args.steps)
        plain_value = ncell * args.steps / tp
        pprof = sim.profile(5)
        sim.destroy()
        log(f"[bench] plain: {tp / args.steps * 1e3:.3f} ms/step, {plain_value:.3e} upd/s, passes {pprof}")
    amr = None
    if args.amr_steps > 0:
        amr = amr_bench(args)
    ens = None
    if args.ens_steps > 0:
        ens = ensemble_bench(args)
    kh = pub = None
    if args.kh_steps > 0:
        kh = kh_bench(args)
        pub = published_bench(args)
    kh_amr = None
    if args.kh_amr_steps > 0:
        kh_amr = kh_amr_bench(args)
    if value is None:
        if amr or ens or kh or kh_amr:
            print(json.dumps({"amr": amr, "ensemble": ens, "kh": kh, "published_grids": pub, "kh_adaptive": kh_amr}))
        return

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (RefHydro restatement) ...")
        cpu = cpu_baseline(pb)
        log(f"[bench] CPU: {cpu['value']:.3e} upd/s on {cpu['cores']} cores")

    out = {
        "metric": "zone-updates/s (256³ + AMR blocks), batched FP32 vs plain, % HBM roofline",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded random-phase solenoidal turbulence, k^-2 on 1<=|k|<=8, rms Mach 1",
        "config": {"workload": f"3D decaying turbulence {n}^3 periodic, gamma 5/3, PLM (MC) + HLLC, "
                               f"MUSCL-Hancock, CFL 0.4", "cells": ncell, "path": "fused one-pass",
                   "tile": tile},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": TRAFFIC_MEASURED.get((kname, n, tuple(tile or ()))),
                     "traffic_source": (f"{TRAFFIC_SOURCE}, {kname} tile {tuple(tile)}"
                                        if (kname, n, tuple(tile or ())) in TRAFFIC_MEASURED else None),
                     "kernel": kname, "kernel_ms": kernel_ms, "bytes_per_update": BYTES_PER_UPDATE,
                     "valu": valu_issue(kname, n, tile, kernel_ms)},
        "per_step": step_roofline(ncell, t / args.steps * 1e3),
        "plain": {"value": plain_value, "ms_per_step": tp / args.steps * 1e3 if tp else None, "passes_ms": pprof},
        "fused_over_plain": value / plain_value if plain_value else None,
        "conserved": {"mass": diag.mass, "e_tot": diag.e_tot},
        "cpu_baseline": cpu,
        "amr": amr,
        "ensemble": ens,
        "kh": kh,
        "published_grids": pub,
        "kh_adaptive": kh_amr,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
