#!/usr/bin/env python3
"""Benchmark of the MARL field path on one MI355X.

Workload (the stencil + light stress grid): 256x256x128 voxels x 12 f32
species, 5% occupancy in smoothed-noise blobs, per-voxel deposit factor in
[0.2, 1], D_s in [0.05, 0.16], 8 masked substeps per tick, then the
Beer-Lambert light scan.  One step = one tick = 8 substeps + 1 light scan on
HBM-resident state.  Default: 125 steps = 1,000 substeps.

Output: one JSON line.  `value` is voxel-species updates per second of the
whole step (G/s).  `roofline` is the masked stencil's algorithmic bytes per
launch (96.125 B per voxel: 12 f32 read + 12 f32 written + the 1-bit mask)
over its average launch time from hipEvents recorded on the library's stream
around the same timed launches.  `cpu_baseline` times the f32 CPU oracle
stencil (OpenMP over z like the reference's rayon loop) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gvoxel-species/s stencil; Mcell-updates/s plain vs surrogate at ≤1e-3 rel err"
HBM_PEAK_GBS = 8000.0
STENCIL_BYTES_PER_VOXEL = 96.125
LIGHT_BYTES_PER_VOXEL = 8.125  # s4 read, intensity written, mask bit


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(field, occ, dims, sim, substeps=400):
    from oracle.binding import Oracle

    orc = Oracle()
    threads = len(os.sched_getaffinity(0))
    f = np.ascontiguousarray(field, np.float32)
    dt_sub = np.float32(sim.dt) / np.float32(sim.diffusion_substeps)
    t0 = time.perf_counter()
    for _ in range(substeps):
        f = orc.substep(f, occ, dims, sim, dt_sub, nthreads=threads)
    dt = time.perf_counter() - t0
    nvs = dims[0] * dims[1] * dims[2] * 12
    return {
        "value": round(substeps * nvs / dt / 1e9, 4),
        "unit": "Gvoxel-species/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{substeps} masked substeps of the 256x256x128x12 stress grid, f32 C oracle, "
                  f"OpenMP over z-layers, dense bool occupancy as the reference's ({dt:.2f} s)",
        "cpu_model": cpu_model(),
    }


def _records_equal(a, b):
    """Structured arrays equal bit for bit in every named field (padding and
    reserved fields, whose names start with '_', are not compared)."""
    if a.dtype.names is None:
        return a.shape == b.shape and a.tobytes() == b.tobytes()
    return all(_records_equal(a[k], b[k]) for k in a.dtype.names if not k.startswith("_"))


def cpu_model() -> str:
    """The host CPU's model name (the `lscpu` "Model name" field, read from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


# The oracle's cell pass and light scan look occupancy up in dense arrays (a
# voxel -> cell index map and a byte mask) where the reference probes its
# HashMap<[u16; 3], usize> (crates/marl-field/src/light.rs:44,
# crates/marl-sim/src/spatial.rs:18-43); the CPU figures below are therefore
# those of a dense-array variant, faster than the reference's own lookups.
DENSE_NOTE = "dense-array occupancy lookups in place of the reference's HashMap probes"

KINETICS_BYTES_PER_CELL = 1350.0  # state r/w + ruleset + neighbour gather/RMW at 12% occupancy


def kinetics_bench(pkg, ticks, with_cpu, reps=5):
    """Plain per-cell kinetics at scale: 256x256x128 with 1,000,000 live cells.

    Throughput: `reps` times, the initial state is uploaded and one full tick is
    run (boundary, 10 substeps, light, ordered cell pass over all 1M cells,
    fate with mutation, compaction); hipEvents split the tick into phases and
    the kinetics launches give the plain-kinetics rate.  Then `ticks` further
    ticks of the trajectory are run for context (the random-reaction population
    shrinks as cells starve)."""
    from marl_ochre_amd import scenarios

    dims = (256, 256, 128)
    log("generating 1M-cell kinetics workload")
    field, cells, rules, sim = scenarios.kinetics_at_scale(dims, 1_000_000, seed=0)
    n0 = len(cells)
    g = pkg.MarlGpu(dims)
    # the random-access ceiling the cell pass works against: 64-byte sectors at
    # hashed positions of a 1 GiB buffer
    import ctypes
    gms, gby = ctypes.c_float(), ctypes.c_double()
    g._chk(g.lib.marl_gather_probe(g.h, 10, ctypes.byref(gms), ctypes.byref(gby)))
    gather_gbs = 10 * gby.value / (gms.value * 1e-3) / 1e9
    log(f"gather probe: {gather_gbs:.0f} GB/s of random 64-byte sectors")
    # the kinetics autotuner on the resident state, then each block size's pass for the record
    g.upload_field_aos(field)
    g.upload_cells(cells, rules)
    chosen = g.autotune_kinetics(sim)
    by_block = {}
    for b in (64, 128, 256):
        g.kinetics_block = b
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        by_block[str(b)] = round(g.tick(sim, 0, 1)["ms_kinetics"], 4)
    g.kinetics_block = chosen
    runs = []
    for rep in range(reps + 1):
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        g.sync()
        t0 = time.perf_counter()
        st = g.tick(sim, 0, 1)
        wall = time.perf_counter() - t0
        if rep > 0:  # the first is a warm-up
            runs.append((wall, st))
    graph_runs = []
    for rep in range(reps + 2):
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        g.sync()
        t0 = time.perf_counter()
        st = g.tick_graph(sim, 0, 1)
        wall = time.perf_counter() - t0
        if rep > 1:  # the first captures the graph, the second re-captures it with the fitted budget
            graph_runs.append((wall, st))
    for (_, a), (_, b) in zip(runs, graph_runs):
        assert all(a[k] == b[k] for k in ("divisions", "deaths", "births_added", "n_cells")), (a, b)
    first_tick = (g.download_field_aos(), *g.download_cells()) if with_cpu else None
    # run outputs on the 1M-cell state: GPU reductions / packing incl. the readback
    g.upload_field_aos(field)
    g.upload_cells(cells, rules)
    g.sync()
    outputs = {}
    for name, fn in (("cell_summary", g.cell_summary), ("chemistry_profile", g.chemistry_profile),
                     ("viewer_records", g.download_viewer_cells)):
        fn()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        outputs[f"ms_{name}"] = round((time.perf_counter() - t0) * 1e3 / 5, 3)
    c0 = cells["internal"].astype(np.float64)
    t0 = time.perf_counter()
    for k in (0, 5, 6):
        np.cumsum(c0[:, k])  # the reference's sequential FP64 sums (data.rs:213-240), NumPy on one core
    (np.abs(rules["reactions"]["v_max"]) > np.float32(1e-9)).sum()
    np.bincount(cells["pos"][:, 2].astype(np.int64), minlength=dims[2])
    outputs["ms_cell_summary_numpy_host"] = round((time.perf_counter() - t0) * 1e3, 3)
    traj = []
    t0 = time.perf_counter()
    for t in range(ticks):
        n_before = g.cell_count()
        traj.append((n_before, g.tick(sim, 1 + t, 1)))
    traj_wall = time.perf_counter() - t0
    g.close()
    kin_ms = sum(st["ms_kinetics"] for _, st in runs) / reps
    tick_ms = sum(w for w, _ in runs) * 1e3 / reps
    res = {
        "workload": "kinetics at scale: 256x256x128, 1,000,000 live cells placed uniformly, 16 random reactions/cell "
                    "(v_max, k_m log-uniform [1e-3,1], p=0.2 light-catalysed), first full tick with mutation, "
                    f"{reps} repetitions from the same state",
        "cells": n0,
        "Mcell_updates_per_s_kinetics": round(n0 / (kin_ms * 1e-3) / 1e6, 3),
        # ms_kinetics is the pass on the tick's critical path: the levels on the
        # AoS view, which the field phase's last substep writes; the transpose
        # back to the quads runs beside fate
        "ms_kinetics_with_transpose_back": round(
            kin_ms + sum(st["ms_transpose_back"] for _, st in runs) / reps, 3),
        "Mcell_updates_per_s_full_tick": round(n0 / (tick_ms * 1e-3) / 1e6, 3),
        "ms_full_tick": round(tick_ms, 3),
        "ms_full_tick_graph": round(sum(w for w, _ in graph_runs) * 1e3 / reps, 3),
        "ms_full_tick_graph_gpu": round(sum(st["ms_field"] for _, st in graph_runs) / reps, 3),
        # the schedule runs beside the field phase and the transpose back beside fate,
        # so the phases overlap; ms_tick_gpu is the tick's own span
        "ms_breakdown": {k: round(sum(st[k] for _, st in runs) / reps, 3)
                         for k in ("ms_field", "ms_schedule", "ms_kinetics", "ms_fate", "ms_transpose_back")},
        "ms_tick_gpu": round(sum(st["ms_tick"] for _, st in runs) / reps, 3),
        "levels": runs[-1][1]["levels"],
        "kinetics_autotune": {"block": chosen, "ms_kinetics_by_block": by_block},
        "events_first_tick": {k: runs[-1][1][k] for k in ("divisions", "deaths", "births_added", "n_cells")},
        "roofline_kinetics": {
            "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved": round(KINETICS_BYTES_PER_CELL * n0 / (kin_ms * 1e-3) / 1e9, 1),
            "bytes_per_cell": KINETICS_BYTES_PER_CELL,
            # what scattered 64-byte reads reach on this box (the pass reads scattered records)
            "random_sector_ceiling_GBs": round(gather_gbs, 1),
        },
        "trajectory": {"ticks": ticks, "ms_per_tick": round(traj_wall * 1e3 / max(ticks, 1), 3),
                       "cells_after": traj[-1][1]["n_cells"] if traj else n0},
        "outputs": outputs,
    }
    res["roofline_kinetics"]["frac"] = round(res["roofline_kinetics"]["achieved"] / HBM_PEAK_GBS, 4)
    if with_cpu:
        from oracle.binding import Oracle

        orc = Oracle()
        threads = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        f_o, _, c_o, r_o, _, _ = orc.tick(dims, sim, 0, field, cells, rules, nthreads=threads)
        dt = time.perf_counter() - t0
        f_g, c_g, r_g = first_tick
        same = bool(np.array_equal(f_g.view(np.uint32), f_o.view(np.uint32)) and len(c_g) == len(c_o)
                    and _records_equal(c_g, c_o) and _records_equal(r_g, r_o))
        res["cpu_baseline_full_tick"] = {
            "value": round(dt * 1e3, 1), "unit": "ms/tick", "cores": threads, "kind": "port",
            "sample": "the first full tick of the 1M-cell workload, f32 C oracle (diffusion OpenMP over z-layers; "
                      f"light, cell pass and fate on one thread, as the reference; {DENSE_NOTE})",
            "cpu_model": cpu_model(),
            "gpu_tick_bit_identical": same}
        res["speedup_full_tick_vs_cpu"] = round(dt * 1e3 / tick_ms, 1)
        nv = dims[0] * dims[1] * dims[2]
        sample = 200_000
        light = np.ones(nv, np.float32)
        t0 = time.perf_counter()
        orc.cell_pass(dims, sim, field, light, cells[:sample], rules[:sample])
        dt = time.perf_counter() - t0
        res["cpu_baseline_cell_pass"] = {
            "value": round(sample / dt / 1e6, 4), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"sequential cell pass (env gather, CellState::tick, delta scatter) over the first {sample} "
                      f"cells of the 1M-cell workload, one thread as in the reference, {DENSE_NOTE} ({dt:.2f} s)",
            "cpu_model": cpu_model()}
    log("kinetics:", json.dumps(res))
    return res


def column_bench(pkg, ticks=200, cpu_ticks=None, with_cpu=True, modes=("graph", "eager")):
    """The reference's default column run: 128x128x64, ~4,096 cells seeded in
    three depth bands (one metabolism each), internals U[0,0.1], 16 random
    reactions (v_max log-U[1e-3,1e-1], k_m log-U[1e-2,1]), linear-gradient
    field, 10 substeps, `ticks` whole ticks from the uploaded state (the first
    tick captures the graph).  The same run with the starters' own states and
    reactions (a population that lives) is timed too.  CPU: the f32 oracle's
    full tick (diffusion OpenMP over z-layers, light and the cell pass on one
    thread, as the reference) over the first `cpu_ticks` ticks (default: all),
    whose end state is compared with the GPU's after the same ticks."""
    from marl_ochre_amd import scenarios

    cpu_ticks = ticks if cpu_ticks is None else cpu_ticks
    dims = (128, 128, 64)
    res = {"workload": "default column 128x128x64, ~4,096 cells in bands [4,12),[24,40),[52,60), internals U[0,0.1], "
                       "16 random reactions (v_max log-U[1e-3,1e-1], k_m log-U[1e-2,1]), linear gradients, "
                       f"10 substeps, {ticks} ticks"}
    for name, starter in (("random_reactions", False), ("starter_states", True)):
        field, cells, rules, sim = scenarios.winogradsky_column(dims, starter_state=starter)
        row = {"cells_start": len(cells)}
        for mode in modes:
            g = pkg.MarlGpu(dims)
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
            g.sync()
            step = g.tick_graph if mode == "graph" else g.tick
            t0 = time.perf_counter()
            step(sim, 0, ticks)
            g.sync()
            row[f"ms_per_tick_{mode}"] = round((time.perf_counter() - t0) * 1e3 / ticks, 4)
            row["cells_end"] = g.cell_count()
            g.close()
        ref_mode = modes[0]
        row[f"Gvoxel_species_substeps_per_s_{ref_mode}"] = round(
            dims[0] * dims[1] * dims[2] * 12 * sim.diffusion_substeps / (row[f"ms_per_tick_{ref_mode}"] * 1e-3) / 1e9,
            1)
        res[name] = row
    if with_cpu:
        from oracle.binding import Oracle

        orc = Oracle()
        threads = len(os.sched_getaffinity(0))
        field, cells, rules, sim = scenarios.winogradsky_column(dims)
        f, c, r = field, cells, rules
        t0 = time.perf_counter()
        for t in range(cpu_ticks):
            f, _, c, r, _, _ = orc.tick(dims, sim, t, f, c, r, nthreads=threads)
        dt = time.perf_counter() - t0
        g = pkg.MarlGpu(dims)
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        (g.tick_graph if modes[0] == "graph" else g.tick)(sim, 0, cpu_ticks)
        c_g, _ = g.download_cells()
        same = bool(np.array_equal(g.download_field_aos().view(np.uint32), f.view(np.uint32))
                    and len(c_g) == len(c)
                    and np.array_equal(c_g["internal"].view(np.uint32), c["internal"].view(np.uint32))
                    and np.array_equal(c_g["pos"], c["pos"]))
        g.close()
        res["cpu_baseline"] = {
            "value": round(dt * 1e3 / cpu_ticks, 3), "unit": "ms/tick", "cores": threads, "kind": "port",
            "sample": f"the first {cpu_ticks} ticks of the random-reaction run, f32 C oracle, {DENSE_NOTE} "
                      f"({dt:.2f} s)",
            "cpu_model": cpu_model()}
        best = min(modes, key=lambda m: res["random_reactions"][f"ms_per_tick_{m}"])
        res["speedup_vs_cpu"] = round(res["cpu_baseline"]["value"] / res["random_reactions"][f"ms_per_tick_{best}"], 1)
        res["speedup_vs_cpu_mode"] = best
        res["gpu_bit_identical_to_oracle"] = {"tick": cpu_ticks, "field_and_cells": same}
    log("column:", json.dumps(res))
    return res


def dropin_bench(pkg, reps=10):
    """The drop-in boundary (GpuFieldDiffuser::diffuse_tick_with_cells,
    crates/marl-gpu/src/field_diffusion.rs:158-292) at the reference's default
    grid, 128x128x64, 10 substeps: host AoS field and bool occupancy in, field
    out, so each call moves 50 MB each way over PCIe.  Beside it the same
    substeps on the resident field, and the f32 CPU stencil (OpenMP)."""
    from oracle.binding import Oracle

    dims = (128, 128, 64)
    rng = np.random.default_rng(1)
    field = rng.random(dims[0] * dims[1] * dims[2] * 12).astype(np.float32)
    occ = (rng.random(dims[0] * dims[1] * dims[2]) < 0.02).astype(np.uint8)
    sim = pkg.SimulationConfig()
    d = pkg.GpuFieldDiffuser.new(dims)
    f = field.copy()
    d.diffuse_tick_with_cells(f, occ, sim)
    t0 = time.perf_counter()
    for _ in range(reps):
        d.diffuse_tick_with_cells(f, occ, sim)
    call_ms = (time.perf_counter() - t0) * 1e3 / reps
    d.close()
    g = pkg.MarlGpu(dims)
    g.upload_field_aos(field)
    g.upload_occupancy(occ)
    g.time_diffuse(sim, 10)
    resident_ms = g.time_diffuse(sim, 10 * reps) / reps
    g.close()
    orc = Oracle()
    nthreads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    orc.diffuse_tick(field, occ, dims, sim, nthreads=nthreads)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    res = {"workload": "GpuFieldDiffuser drop-in at 128x128x64, 10 substeps, 2% occupancy; host buffers in and out",
           "ms_per_call_pcie_inclusive": round(call_ms, 3), "ms_resident_10_substeps": round(resident_ms, 3),
           "ms_cpu_oracle_10_substeps": round(cpu_ms, 3), "cpu_threads": nthreads,
           "bytes_over_pcie_per_call": 2 * field.nbytes + occ.nbytes}
    log("drop-in:", json.dumps(res))
    return res


def drift_report(pkg, n_ticks, nthreads, tile=-1):
    """Long-run accuracy on the default 128x128x64 column (~4,096 cells).

    (1) Full ticks on the GPU and in the FP32 oracle in lockstep for n_ticks:
        field, light and every cell field compared at checkpoints.
    (2) Field-only ticks (boundary sources + 10 substeps, occupancy of the
        initial cells held fixed) on the GPU in FP32 against the oracle's FP64
        statement: relative L2 / max errors and per-species mass at checkpoints."""
    from marl_ochre_amd import scenarios
    from oracle.binding import Oracle

    orc = Oracle()
    dims = (128, 128, 64)
    field, cells, rules, sim = scenarios.winogradsky_column(dims, starter_state=True)
    checks = sorted({c for c in (1, 10, 100, 1000, n_ticks) if c <= n_ticks})
    out = {"workload": "default column 128x128x64 with the starters' own states, %d cells at start, 10 substeps/tick, "
                       "GPU ticks through the captured graph" % len(cells), "ticks": n_ticks,
           "stencil_tile": tile if tile >= 0 else "default"}
    # (1) full ticks, FP32 GPU vs FP32 oracle
    g = pkg.MarlGpu(dims)
    if tile >= 0:
        g.tile = tile
    g.upload_field_aos(field)
    g.upload_cells(cells, rules)
    f_o, c_o, r_o = field.copy(), cells.copy(), rules.copy()
    rows = []
    t0 = time.perf_counter()
    for t in range(n_ticks):
        g.tick_graph(sim, t, 1)
        f_o, l_o, c_o, r_o, _, _ = orc.tick(dims, sim, t, f_o, c_o, r_o, nthreads=nthreads)
        if t + 1 in checks:
            f_g = g.download_field_aos()
            l_g = g.download_light()
            c_g, r_g = g.download_cells()
            same_cells = len(c_g) == len(c_o) and all(
                np.array_equal(c_g[k], c_o[k]) for k in ("pos", "lineage_id", "age", "prep_remaining", "quiescent"))
            same_rules = len(r_g) == len(r_o) and all(
                np.array_equal(r_g["reactions"][k], r_o["reactions"][k]) for k in r_g["reactions"].dtype.names) and \
                all(np.array_equal(r_g[k], r_o[k]) for k in ("division_energy", "death_energy", "mutation_rate"))
            rows.append({"tick": t + 1, "cells": len(c_g), "field_max_abs_diff": float(np.abs(f_g - f_o).max()),
                         "light_max_abs_diff": float(np.abs(l_g - l_o).max()),
                         "internal_max_abs_diff": float(np.abs(c_g["internal"] - c_o["internal"]).max())
                         if same_cells else None,
                         "bit_identical": bool(same_cells and same_rules
                                               and np.array_equal(f_g.view(np.uint32), f_o.view(np.uint32))
                                               and np.array_equal(l_g.view(np.uint32), l_o.view(np.uint32))
                                               and np.array_equal(c_g["internal"].view(np.uint32),
                                                                  c_o["internal"].view(np.uint32)))})
            log("drift fp32:", json.dumps(rows[-1]), f"({time.perf_counter() - t0:.1f} s)")
        elif (t + 1) % 100 == 0:
            log(f"drift fp32: tick {t + 1}")
    g.close()
    out["fp32_full_tick_vs_oracle"] = rows
    # (2) field-only ticks, FP32 GPU vs FP64 oracle
    occ = np.zeros(dims[0] * dims[1] * dims[2], np.uint8)
    p = cells["pos"].astype(np.int64)
    occ[(p[:, 2] * dims[1] + p[:, 1]) * dims[0] + p[:, 0]] = 1
    g = pkg.MarlGpu(dims)
    if tile >= 0:
        g.tile = tile
    g.upload_field_aos(field)
    g.upload_occupancy(occ)
    f64 = field.astype(np.float64)
    rows, done = [], 0
    for c in checks:
        for _ in range(c - done):
            g.apply_boundary_sources(sim)
            g.diffuse(sim)
        orc.field_ticks_f64(f64, occ, dims, sim, c - done)
        done = c
        f32 = g.download_field_aos().astype(np.float64)
        d = f32 - f64
        big = np.abs(f64) > 1e-3
        m32 = f32.reshape(-1, 12).sum(0)
        m64 = f64.reshape(-1, 12).sum(0)
        live = m64 > 1e-9
        rows.append({"tick": c, "rel_l2": float(np.linalg.norm(d) / np.linalg.norm(f64)),
                     "max_abs": float(np.abs(d).max()),
                     "max_rel_where_gt_1e-3": float((np.abs(d[big]) / np.abs(f64[big])).max()),
                     "mass_max_rel_per_species": float((np.abs(m32 - m64)[live] / m64[live]).max())})
        log("drift fp64:", json.dumps(rows[-1]))
    g.close()
    out["fp32_field_vs_fp64_oracle"] = rows
    return out


def ensemble_bench(pkg, ticks=5, warm=2):
    """32 independent columns, x = y in {64, 96, 128, 192}, z = x/2, own seeds,
    starter populations of log-uniform size in [1k, 200k] (capped at 15 % of
    the voxels).  Field phase: one batched substep over all columns against
    the sum of per-column substeps.  Whole ticks: the ensemble's hipGraph
    against ticking the columns one after another (host-driven and graphed),
    each timing ticks warm..warm+ticks-1 of the same trajectories."""
    from marl_ochre_amd import scenarios

    log("generating 32-column ensemble")
    cols = scenarios.ensemble_columns(32, seed=0)
    sim = pkg.SimulationConfig()
    seeds = np.array([c[4] for c in cols], np.uint64)
    nvox = sum(d[0] * d[1] * d[2] for d, *_ in cols)
    ncells = sum(len(c[2]) for c in cols)

    blocks = []   # each column's plain-pass block size, as the ensemble autotuner picked it

    def fresh():
        gs = [pkg.MarlGpu(d) for d, *_ in cols]
        for g, (dims, field, cells, rules, _) in zip(gs, cols):
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
        for g, b in zip(gs, blocks):
            g.kinetics_block = b
        return gs

    gs = fresh()
    res = {"workload": "32 columns x=y in {64,96,128,192}, z=x/2, own seeds, starter populations log-U[1k,200k] "
                       "(<=15% of voxels), 10 substeps/tick",
           "columns": len(cols), "voxels": nvox, "cells": ncells}
    try:
        with pkg.MarlEnsemble(gs) as ens:
            tile = ens.autotune(sim)
            blocks[:] = [g.kinetics_block for g in gs]
            shape = ens.launch_shape()
            shape["kinetics_blocks"] = {str(b): blocks.count(b) for b in sorted(set(blocks))}
            shape["chunks_by_depth"] = {str(d[2]): z for (d, *_), z in sorted(zip(cols, ens.chunks()),
                                                                                key=lambda t: t[0][0][2])}
            ms_sub = ens.time_field(sim, 20)
            ms_phase = ens.time_field_phase(sim, 5)
            for g, (dims, field, cells, rules, _) in zip(gs, cols):
                g.upload_field_aos(field)
                g.upload_cells(cells, rules)
            # the first ticks capture the graph and fit the schedule budgets;
            # every mode below times the same later ticks of the same trajectories
            ens.tick(sim, seeds, 0, warm)
            t0 = time.perf_counter()
            _, gpu_ms, hf = ens.tick(sim, seeds, warm, ticks)
            ens_wall = (time.perf_counter() - t0) * 1e3 / ticks
        per_col_sub = 0.0
        for g in gs:
            g.tile = tile
            per_col_sub += g.time_diffuse(sim, 20) / 20
    finally:
        for g in gs:
            g.close()
    seq = {}
    for mode in ("eager", "graph"):
        gs = fresh()
        try:
            step = {"eager": lambda g, s2, t: g.tick(s2, t, 1), "graph": lambda g, s2, t: g.tick_graph(s2, t, 1)}[mode]
            for t in range(warm):
                for g, c in zip(gs, cols):
                    step(g, sim.replace(rng_seed=c[4]), t)
            for g in gs:
                g.sync()
            t0 = time.perf_counter()
            for t in range(warm, warm + ticks):
                for g, c in zip(gs, cols):
                    step(g, sim.replace(rng_seed=c[4]), t)
            seq[mode] = (time.perf_counter() - t0) * 1e3 / ticks
        finally:
            for g in gs:
                g.close()
    alg = 96.125 * nvox
    res.update({
        "tile": tile,
        "launch": shape,
        "batched_substep_us": round(ms_sub * 1e3, 1),
        "per_column_substeps_us": round(per_col_sub * 1e3, 1),
        "batched_substep_Gvoxel_species_per_s": round(nvox * 12 / (ms_sub * 1e-3) / 1e9, 1),
        "batched_substep_hbm_frac": round(alg / (ms_sub * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        # a tick's field phase: 10 substeps and the light scan (4 B/voxel write when fused into the
        # last substep, 8.125 B/voxel as its own launch)
        "batched_field_phase_us": round(ms_phase * 1e3, 1),
        "batched_field_phase_hbm_frac": round(
            (alg * sim.diffusion_substeps + (4.0 if shape["whole_columns"] else LIGHT_BYTES_PER_VOXEL) * nvox)
            / (ms_phase * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "ms_per_tick_ensemble_graph": round(ens_wall, 3),
        "ms_per_tick_ensemble_graph_gpu": round(gpu_ms / ticks, 3),
        "ms_per_tick_columns_one_by_one_eager": round(seq["eager"], 3),
        "ms_per_tick_columns_one_by_one_graph": round(seq["graph"], 3),
        "speedup_ensemble_vs_eager": round(seq["eager"] / ens_wall, 2),
        "host_finished_column_ticks": hf,
    })
    log("ensemble:", json.dumps(res))
    return res


SURR_FLOP_PER_CELL = 2.0 * (80 * 256 + 256 * 256 * 2 + 256 * 28)  # forward multiply-adds x 2
# a training pair: the forward, the input gradients of layers 4..2 and every weight gradient
SURR_TRAIN_FLOP_PER_PAIR = SURR_FLOP_PER_CELL + 2.0 * (28 * 256 + 256 * 256 * 2) + SURR_FLOP_PER_CELL
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _shadow_summary(sh):
    """Shadow statistics: the within-1e-3 share (all of internals 0-14, and per
    species), the median bin of the per-cell max relative error, fate agreement
    and the fate event rates under the exact update and under the prediction."""
    n = max(sh["shadow"], 1)
    hist = sh["err_hist"]
    med = next((b for b in range(24) if sum(hist[:b + 1]) * 2 >= n), 23)
    ev = ("none", "division", "death", "quiescence")
    return {
        "cells": sh["shadow"],
        "within_1e-3_rel": round(sh["shadow_within_1e3"] / n, 5),
        "within_1e-3_rel_per_species": [round(w / n, 4) for w in sh["shadow_within_1e3_species"][:15]],
        "median_max_rel_err_bin": [10 ** (med / 2 - 9), 10 ** ((med + 1) / 2 - 9)],
        "max_rel_err": sh["shadow_max_rel_err"],
        "fate_agreement": round(sh["shadow_fate_agree"] / n, 5),
        "fate_rate_exact": {e: round(v / n, 5) for e, v in zip(ev, sh["fate_exact"])},
        "fate_rate_surrogate": {e: round(v / n, 5) for e, v in zip(ev, sh["fate_surrogate"])},
        "err_hist_half_decades_from_1e-9": hist,
    }


XGMI_LINK_GBS = 153.0  # one MI355X xGMI link, per direction


def slab_bench(pkg, slabs=8, reps=4):
    """One GPU's share of a single large grid split into `slabs` y-slabs (the
    split that keeps every light column on one GPU): the stress grid's
    256 x 256/slabs x 128 slab with its own autotuned stencil tile, and the
    1M-cell workload's share (1,000,000/slabs cells at the same density) on
    that slab.  No halo exchange runs (one GPU); the per-substep halo plane a
    neighbour would send (256 x 128 voxels x 12 species x 4 B per face) is
    priced at one xGMI link and set against the slab's substep time."""
    from marl_ochre_amd import scenarios

    dims = (256, 256 // slabs, 128)
    nvox = dims[0] * dims[1] * dims[2]
    field, occ, sim = scenarios.stencil_stress(dims, seed=0)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_occupancy(occ)
        tile = g.autotune(sim)
        g.upload_field_aos(field)
        g.time_field_phase(sim, 3)
        n = 40
        phase_ms = g.time_field_phase(sim, n) / n
        sub_ms = g.time_diffuse(sim, 10 * sim.diffusion_substeps) / (10 * sim.diffusion_substeps)
        zc = g.tile_shape(tile)[1]
    kfield, kcells, krules, ksim = scenarios.kinetics_at_scale(dims, 1_000_000 // slabs, seed=0)
    with pkg.MarlGpu(dims) as g:
        g.kinetics_block = 64
        ticks = []
        for rep in range(reps + 1):
            g.upload_field_aos(kfield)
            g.upload_cells(kcells, krules)
            st = g.tick(ksim, 0, 1)
            if rep:
                ticks.append(st)
    med = lambda k: float(np.median([st[k] for st in ticks]))
    halo_bytes = dims[0] * dims[2] * 12 * 4
    halo_us = halo_bytes / (XGMI_LINK_GBS * 1e9) * 1e6
    return {
        "workload": f"one of {slabs} y-slabs: {dims[0]}x{dims[1]}x{dims[2]} of the stress grid, and "
                    f"{len(kcells):,} cells of the 1M-cell workload on it; no halo exchange (one GPU)",
        "slab": list(dims), "stencil_tile": tile, "z_chunk": zc,
        "substep_us": round(sub_ms * 1e3, 2),
        "substep_hbm_frac": round(STENCIL_BYTES_PER_VOXEL * nvox / (sub_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "field_phase_us": round(phase_ms * 1e3, 1),
        "field_phase_Gvoxel_species_per_s": round(nvox * 12 * sim.diffusion_substeps / (phase_ms * 1e-3) / 1e9, 1),
        "halo_plane_bytes_per_face": halo_bytes,
        "halo_plane_us_per_face_one_xgmi_link": round(halo_us, 2),
        "halo_over_substep": round(halo_us / (sub_ms * 1e3), 3),
        "cells": len(kcells),
        "ms_tick": round(med("ms_tick"), 4), "ms_kinetics": round(med("ms_kinetics"), 4),
        "ms_field": round(med("ms_field"), 4), "ms_fate": round(med("ms_fate"), 4), "levels": int(med("levels")),
    }


def surrogate_bench(pkg, train_steps, harvest_ticks, reps=3, pair_target=10_000_000, harvest_seeds=4):
    """The learned kinetics surrogate on the 1M-cell workload, every cell with
    its own random reaction wiring.

    Pairs (explicit-step features, exact-minus-explicit residuals) are harvested
    from plain ticks of the same workload drawn with seeds 1, 2, ... (each up
    to `harvest_ticks` ticks of its trajectory) until `pair_target` pairs.  The
    evaluated tick is the first tick of seed 0, whose states and wirings no
    training pair saw.  For each training-set size (1M, 3M and all pairs)
    the network is trained from scratch on the GPU (batch 8192) and that
    tick is run in shadow mode (plain authoritative, every eligible cell also
    predicted) for its error distribution; the explicit step alone (residual 0)
    is the curve's zero point.  With the network trained on all pairs the
    tick is timed `reps` times in plain and in surrogate mode (kinetics
    launches, hipEvents), and the fused inference kernel alone on 1M rows."""
    from marl_ochre_amd import scenarios

    A = pkg._abi
    dims = (256, 256, 128)
    g = pkg.MarlGpu(dims)
    try:
        g.harvest_begin(pair_target, 1.0)   # pairs beyond the capacity are not kept
        harvest, harvest_s, ticks = [], 0.0, 0
        for hs in range(1, harvest_seeds + 1):
            log(f"generating 1M-cell surrogate workload, seed {hs} (per-cell wiring)")
            hfield, hcells, hrules, hsim = scenarios.kinetics_at_scale(dims, 1_000_000, seed=hs)
            g.upload_field_aos(hfield)
            g.upload_cells(hcells, hrules)
            del hfield, hcells, hrules
            before, t = g.harvest_count(), 0
            t0 = time.perf_counter()
            while t < harvest_ticks and g.harvest_count() < pair_target:
                g.tick(hsim, t, 1)
                t += 1
            harvest_s += time.perf_counter() - t0
            ticks += t
            harvest.append({"seed": hs, "ticks": t, "pairs": g.harvest_count() - before})
            if g.harvest_count() >= pair_target:
                break
        n_pairs = g.harvest_end()
        log(f"harvested {n_pairs} pairs in {ticks} ticks ({harvest_s:.2f} s): {harvest}")
        log("generating the evaluated 1M-cell workload, seed 0")
        field, cells, rules, sim = scenarios.kinetics_at_scale(dims, 1_000_000, seed=0)
        n0 = len(cells)

        def first_tick(mode, shadow_rate=0.0):
            out = []
            for r in range(reps + 1 if mode != A.KINETICS_SHADOW else 1):
                g.set_kinetics_mode(A.KINETICS_PLAIN)
                g.upload_field_aos(field)
                g.upload_cells(cells, rules)
                if mode != A.KINETICS_PLAIN:
                    g.surrogate_configure(zmax=6.0, energy_margin=0.05, shadow_rate=shadow_rate)
                    g.set_kinetics_mode(mode)
                g.surrogate_reset_stats()
                st = g.tick(sim, 0, 1)
                if r > 0 or mode == A.KINETICS_SHADOW:
                    out.append(st)
            stats = g.surrogate_stats()
            g.set_kinetics_mode(A.KINETICS_PLAIN)
            return out, stats

        curve = []
        g.surrogate_set_dead_zone(0.2)   # the shipped operating point
        for size in [m for m in (1_000_000, 3_000_000) if m < n_pairs] + [n_pairs]:
            rep = g.surrogate_train(steps=train_steps, batch=8192, lr=2e-3, lr_final=1e-5, seed=1, reinit=True,
                                    max_pairs=size)
            _, sh = first_tick(A.KINETICS_SHADOW)
            curve.append({"train_pairs": rep["pairs"], "val_loss": round(rep["val_loss"], 5),
                          "train_ms": round(rep["ms"], 1), "shadow": _shadow_summary(sh)})
            log("surrogate curve point:", json.dumps(curve[-1]["shadow"])[:300])
        # the explicit step alone: last layer and target means zeroed, so the residual is 0
        w, norm = g.surrogate_get_params()
        w0, nm0 = w.copy(), norm.copy()
        w0[-(256 * 28 + 28):] = 0.0
        nm0[160:188] = 0.0
        g.surrogate_set_params(w0, nm0)
        _, sh0 = first_tick(A.KINETICS_SHADOW)
        g.surrogate_set_params(w, norm)
        # the network trained on all pairs with corrections below a dead zone dropped
        dead_zone = []
        for dz in (0.0, 0.05, 0.1, 0.2, 0.5, 1.0):
            g.surrogate_set_dead_zone(dz)
            _, shz = first_tick(A.KINETICS_SHADOW)
            dead_zone.append({"dead_zone": dz, "shadow": _shadow_summary(shz)})
            log("surrogate dead zone:", json.dumps(dead_zone[-1])[:300])
        g.surrogate_set_dead_zone(0.2)   # the shipped operating point
        plain, _ = first_tick(A.KINETICS_PLAIN)
        surr, surr_stats = first_tick(A.KINETICS_SURROGATE)
        infer_ms = g.surrogate_time_infer(n0, 10)
        import ctypes

        mms, mfl = ctypes.c_float(), ctypes.c_double()
        g._chk(g.lib.marl_mfma_probe(g.h, 5, ctypes.byref(mms), ctypes.byref(mfl)))
        mfma_meas = 5 * mfl.value / (mms.value * 1e-3) / 1e12
        log(f"mfma probe: {mfma_meas:.0f} TFLOP/s dense BF16")
    finally:
        g.close()
    kin_plain = sum(st["ms_kinetics"] for st in plain) / reps
    kin_surr = sum(st["ms_kinetics"] for st in surr) / reps
    res = {
        "workload": "surrogate: 256x256x128, 1,000,000 cells, 16 random reactions per cell with its own wiring, "
                    "v_max/k_m log-uniform [1e-3,1], p=0.2 light-catalysed; first full tick",
        "pairs_harvested": n_pairs, "harvest_ticks": ticks, "harvest": harvest,
        "evaluated": "first tick of seed 0; training pairs come from seeds 1+",
        "train": {k: curve[-1][k] for k in ("train_pairs", "val_loss", "train_ms")},
        "train_pairs_per_s": round(train_steps * 8192 / (curve[-1]["train_ms"] * 1e-3), 0),
        "train_TFLOPs": round(train_steps * 8192 * SURR_TRAIN_FLOP_PER_PAIR / (curve[-1]["train_ms"] * 1e-3) / 1e12, 2),
        "train_mfma_frac": round(train_steps * 8192 * SURR_TRAIN_FLOP_PER_PAIR / (curve[-1]["train_ms"] * 1e-3) / 1e12
                                 / MFMA_BF16_PEAK_TFLOPS, 5),
        "error_vs_training_pairs": [{"train_pairs": 0, "note": "explicit step alone", "shadow": _shadow_summary(sh0)}]
        + curve,
        "error_vs_dead_zone": dead_zone,
        "Mcell_updates_per_s_plain": round(n0 / (kin_plain * 1e-3) / 1e6, 3),
        "Mcell_updates_per_s_surrogate": round(n0 / (kin_surr * 1e-3) / 1e6, 3),
        "speedup_vs_plain": round(kin_plain / kin_surr, 3),
        "ms_kinetics_plain": round(kin_plain, 3), "ms_kinetics_surrogate": round(kin_surr, 3),
        "surrogate_share": round(surr_stats["surrogate"] / max(surr_stats["cells"], 1), 4),
        "fallbacks": {k: surr_stats[k] for k in ("fallback_topology", "fallback_ood", "fallback_margin")},
        "fate_events_plain": {k: plain[-1][k] for k in ("divisions", "deaths", "births_added", "n_cells")},
        "fate_events_surrogate": {k: surr[-1][k] for k in ("divisions", "deaths", "births_added", "n_cells")},
        "infer_kernel": {"rows": n0, "ms": round(infer_ms, 4),
                         "TFLOPs": round(SURR_FLOP_PER_CELL * n0 / (infer_ms * 1e-3) / 1e12, 1),
                         "mfma_frac": round(SURR_FLOP_PER_CELL * n0 / (infer_ms * 1e-3) / 1e12
                                            / MFMA_BF16_PEAK_TFLOPS, 4),
                         "mfma_frac_of_measured": round(SURR_FLOP_PER_CELL * n0 / (infer_ms * 1e-3) / 1e12
                                                        / mfma_meas, 4),
                         "flop_per_cell": SURR_FLOP_PER_CELL, "peak_TFLOPs": MFMA_BF16_PEAK_TFLOPS},
        "mfma_probe": {"TFLOPs": round(mfma_meas, 1), "frac_of_vendor_peak": round(mfma_meas / MFMA_BF16_PEAK_TFLOPS, 4),
                       "kernel": "register-only v_mfma_f32_32x32x16_bf16 chains, 4 per wave, 4 blocks of 256 per CU"},
    }
    log("surrogate:", json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=125)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (host buffers) leg")
    ap.add_argument("--tile", type=int, default=-1, help="use this stencil tile id instead of autotuning")
    ap.add_argument("--kinetics-ticks", type=int, default=100,
                    help="full ticks of the 1M-cell kinetics workload (0 = skip)")
    ap.add_argument("--column-ticks", type=int, default=200,
                    help="ticks of the default 128x128x64 column run (0 = skip)")
    ap.add_argument("--column-modes", default="graph,eager",
                    help="tick paths of the column run: graph (marl_tick_graph), eager (marl_tick)")
    ap.add_argument("--slabs", type=int, default=8,
                    help="time one GPU's y-slab of the stress grid and 1M-cell workload split this many ways (<= 1 = skip)")
    ap.add_argument("--ensemble-ticks", type=int, default=5,
                    help="ticks of the 32-column ensemble leg (0 = skip)")
    ap.add_argument("--drift-ticks", type=int, default=0,
                    help="long-run accuracy report over this many ticks of the default column (0 = skip)")
    ap.add_argument("--surrogate-steps", type=int, default=20000,
                    help="training steps of the surrogate leg (0 = skip)")
    ap.add_argument("--harvest-ticks", type=int, default=20,
                    help="plain ticks per harvest seed whose pairs train the surrogate (up to 10M pairs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (ensembles would shard "
              "by column, a single grid by y-slabs with one halo-plane exchange per substep)", file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package

    pkg = load_package()
    from marl_ochre_amd import scenarios  # noqa: E402  (registered by load_package)

    dims = (256, 256, 128)
    log("generating stress grid", dims)
    field, occ, sim = scenarios.stencil_stress(dims, seed=0)
    nvox = dims[0] * dims[1] * dims[2]

    g = pkg.MarlGpu(dims)
    g.upload_field_aos(field)
    g.upload_occupancy(occ)
    if args.tile >= 0:
        g.tile = args.tile
    tile = g.tile if (args.no_autotune or args.tile >= 0) else g.autotune(sim)
    ry, zc, pd = g.tile_shape(tile)
    sg, nw = g.tile_split(tile)
    log(f"stencil tile {tile}: rows/thread={ry} z-chunk={zc} prefetch={pd} species/thread={sg} waves/block={nw}")

    copy_ms, copy_bytes = 0.0, 0.0
    import ctypes

    ms = ctypes.c_float()
    by = ctypes.c_double()
    g._chk(g.lib.marl_copy_probe(g.h, 20, ctypes.byref(ms), ctypes.byref(by)))
    copy_gbs = 20 * by.value / (ms.value * 1e-3) / 1e9
    log(f"copy probe: {copy_gbs:.0f} GB/s")

    # one step = one tick's field phase as the tick runs it: 8 substeps, the
    # light scan fused into the last when the tile's z-chunk spans the column
    fused = zc >= dims[2] and g.tile_substeps(tile) == 1
    for _ in range(args.warmup):
        g.time_field_phase(sim, 1)
    g.sync()

    # ticks run back to back, five per timed call (as a run would), so the
    # host round trip of each call is spread over five ticks
    phase_ms = 0.0
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(5, args.steps - done)
        phase_ms += g.time_field_phase(sim, k)
        done += k
    g.sync()
    wall = time.perf_counter() - t0

    n_sub = args.steps * sim.diffusion_substeps
    ms_per_step = wall * 1e3 / args.steps
    value = n_sub * nvox * 12 / wall / 1e9
    # the dominant kernel is k_stencil: with the light fused, the step is its
    # 8 launches, the last also writing the light (4 B per voxel; its s4 and
    # mask reads are the stencil's own); unfused, the step also holds k_light
    per_launch_s = phase_ms * 1e-3 / n_sub
    light_bytes = (4.0 if fused else LIGHT_BYTES_PER_VOXEL) * nvox
    alg_bytes = STENCIL_BYTES_PER_VOXEL * nvox + light_bytes / sim.diffusion_substeps
    achieved = alg_bytes / per_launch_s / 1e9
    # the separate light kernel and plain substeps, for reference
    light_ms = g.time_light(sim, 20) / 20
    plain_sub_ms = g.time_diffuse(sim, 5 * sim.diffusion_substeps) / (5 * sim.diffusion_substeps)
    log(f"field phase {phase_ms / args.steps * 1e3:.1f} us/tick (light fused: {fused}), "
        f"{per_launch_s * 1e6:.1f} us/launch ({achieved:.0f} GB/s algorithmic); plain substep "
        f"{plain_sub_ms * 1e3:.1f} us, separate light {light_ms * 1e3:.1f} us")

    # HBM bytes per launch from the committed rocprofv3 PMC summary, valid for
    # the tile it was measured on
    traffic = None
    tf = ROOT / "profiles" / "stencil_traffic.json"
    if tf.exists():
        meas = json.loads(tf.read_text())
        ent = meas.get("tiles", {}).get(str(tile))
        if ent and tuple(meas.get("grid", ())) == dims:
            traffic = ent.get("bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxel-species/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "stencil+light stress grid 256x256x128x12, 5% occupancy blobs, deposit factor U[0.2,1], "
                        "8 masked substeps + 1 light scan per step (1 step = 1 tick)",
            "grid": list(dims),
            "substeps_per_step": sim.diffusion_substeps,
            "stencil_tile": {"id": tile, "rows_per_thread": ry, "z_chunk": zc, "prefetch_planes": pd},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_stencil (one masked substep; the last of a tick also writes the light scan)" if fused
                      else "k_stencil (one masked substep) + k_light per tick",
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_us": round(per_launch_s * 1e6, 2),
            "measured_copy_GBs": round(copy_gbs, 1),
            "plain_substep_us": round(plain_sub_ms * 1e3, 2),
            "plain_substep_frac": round(STENCIL_BYTES_PER_VOXEL * nvox / (plain_sub_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "light_fused": fused,
        "light_us_per_tick_separate_kernel": round(light_ms * 1e3, 2),
    }
    g.close()
    if args.kinetics_ticks > 0:
        out["kinetics"] = kinetics_bench(pkg, args.kinetics_ticks, not args.no_cpu)
    if args.column_ticks > 0:
        out["column"] = column_bench(pkg, args.column_ticks, with_cpu=not args.no_cpu,
                                     modes=tuple(m for m in args.column_modes.split(",") if m))
    if args.ensemble_ticks > 0:
        out["ensemble"] = ensemble_bench(pkg, args.ensemble_ticks)
    if not args.no_dropin:
        out["dropin"] = dropin_bench(pkg)
    if args.drift_ticks > 0:
        out["drift"] = drift_report(pkg, args.drift_ticks, min(os.cpu_count() or 1, 16), args.tile)
    if args.slabs > 1:
        out["slab"] = slab_bench(pkg, args.slabs)
    if args.surrogate_steps > 0:
        out["surrogate"] = surrogate_bench(pkg, args.surrogate_steps, args.harvest_ticks)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(field, occ, dims, sim)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
