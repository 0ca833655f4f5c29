"""Literal Python restatements of the reference's fate bookkeeping, mutation
and seeding, written from the reference's source and independent of the
oracle's C code.  Test infrastructure: they pin the oracle functions
orc_tick (fate, birth, death and compaction), orc_mutate and
orc_seed_positions, which no reference output can pin (the reference draws
from an OS-seeded ThreadRng, crates/marl-sim/src/lib.rs:35).

The only primitives taken from the oracle are the counter-based draws
(orc_draw_u32, pinned to numpy's Philox golden vectors), orc_logf (pinned to
libm) and, for the tick, find_empty_neighbor and mutate — each pinned by its
own test.  The stream definitions restated here (uniform f32, bounded range,
Marsaglia polar normal, u64 lineage from two draws) are this project's
replacement of rand 0.9.2 / rand_distr 0.5.1 and are written out from their
definitions:
  u01(u)      = (u >> 8) * 2^-24            (24 bits, exact in f32)
  range(u, k) = (u * k) >> 32               (widening multiply)
  normal(sd)  = 0 + sd * u * sqrt(-2 ln s / s), (u, v) = 2 * u01 - 1 until
                0 < s = u^2 + v^2 < 1       (Marsaglia polar method)
  u64         = lo | hi << 32, lo drawn first
All f32 arithmetic is numpy float32 scalars, one IEEE operation per step.
"""
from __future__ import annotations

import numpy as np

F = np.float32
TAG_SEEDING = 1
TAG_DIVISION = 2
FACE_OFFSETS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]  # spatial.rs:1-7
EVENT_DIVISION, EVENT_DEATH = 1, 2


class Stream:
    """One entity's draw sequence (key (seed, stream), entity, next draw index)."""

    def __init__(self, orc, seed, stream, entity, draws=0):
        self.orc, self.seed, self.stream, self.entity, self.draws = orc, seed, stream, entity, draws

    def u32(self) -> int:
        v = self.orc.draw_u32(self.seed, self.stream, self.entity, self.draws)
        self.draws += 1
        return int(v)

    def u01(self) -> np.float32:
        return F(self.u32() >> 8) * F(2.0 ** -24)

    def range(self, k: int) -> int:
        return (self.u32() * k) >> 32

    def u64(self) -> int:
        lo = self.u32()
        hi = self.u32()
        return lo | (hi << 32)

    def normal(self, sd) -> np.float32:
        while True:
            u = F(2.0) * self.u01() - F(1.0)
            v = F(2.0) * self.u01() - F(1.0)
            s = u * u + v * v
            if F(0.0) < s < F(1.0):
                break
        z = u * np.sqrt((F(-2.0) * self.orc.logf(s)) / s)
        return F(0.0) + F(sd) * z


def _clamp(x, lo, hi):
    """f32::clamp: min when below, max when above."""
    if x < lo:
        x = F(lo)
    if x > hi:
        x = F(hi)
    return F(x)


def mutate(rule, sim, rs: Stream):
    """Ruleset::mutate — crates/marl-cell/src/cell.rs:406-524, statement by
    statement; returns a mutated copy.  `rs` carries the draw position."""
    r = rule.copy()
    rate = F(r["mutation_rate"])
    sd = F(sim.mutation_stddev)
    struct_rate = rate * F(sim.structural_mutation_rate_mult)

    def maybe(v):  # :411-416
        if rs.u01() < rate:
            v = F(v) + rs.normal(sd)
            v = max(v, F(0.0))
        return F(v)

    for k in range(8):  # receptors, :419-426
        rc = r["receptors"][k]
        rc["k_half"] = maybe(rc["k_half"])
        rc["n_hill"] = maybe(rc["n_hill"])
        rc["n_hill"] = _clamp(F(rc["n_hill"]), sim.hill_exponent_clamp_low, sim.hill_exponent_clamp_high)
        rc["gain"] = maybe(rc["gain"])
    for k in range(8):  # transport, :429-438
        t = r["transport"][k]
        t["uptake_rate"] = maybe(t["uptake_rate"])
        t["secrete_rate"] = maybe(t["secrete_rate"])
        if rs.u01() < struct_rate:
            t["ext_species"] = rs.range(12)
        if rs.u01() < struct_rate:
            t["int_species"] = rs.range(16)
    rx = r["reactions"]
    active = [i for i in range(16) if abs(F(rx[i]["v_max"])) > F(sim.active_reaction_threshold)]  # :449-455
    for i in range(16):  # :457-499
        rx[i]["k_m"] = maybe(rx[i]["k_m"])
        rx[i]["v_max"] = maybe(rx[i]["v_max"])
        rx[i]["k_cat"] = maybe(rx[i]["k_cat"])
        for name in ("substrate", "product", "catalyst"):
            if rs.u01() < struct_rate:
                idx = rs.range(max(len(active), 1))
                if idx < len(active):
                    donor = active[idx]
                    if donor != i:
                        rx[i][name] = rx[donor][name]
                else:
                    rx[i][name] = rs.range(16)
    for k in range(8):  # effectors, :502-505
        e = r["effectors"][k]
        e["threshold"] = maybe(e["threshold"])
        e["rate"] = maybe(e["rate"])
    for name in ("division_energy", "death_energy", "quiescence_energy", "division_prep_ticks"):  # :508-512
        r[name] = maybe(r[name])
    r["division_prep_ticks"] = max(F(r["division_prep_ticks"]), F(1.0))  # :513
    r["hgt_propensity"] = maybe(r["hgt_propensity"])  # :514
    if rs.u01() < F(sim.meta_mutation_rate):  # :518-523
        r["mutation_rate"] = F(r["mutation_rate"]) + rs.normal(sd) * F(sim.meta_mutation_rate)
        r["mutation_rate"] = _clamp(F(r["mutation_rate"]), sim.meta_mutation_clamp_low, sim.meta_mutation_clamp_high)
    return r


def _vox(dims, p):
    return (int(p[2]) * dims[1] + int(p[1])) * dims[0] + int(p[0])


def fate_and_compaction(orc, dims, sim, tick, cells, rules, events):
    """crates/marl-sim/src/lib.rs:220-277 line by line, after the cell pass.

    cells/rules: the population after the pass (Vec order), events: the pass's
    CellEvent per cell.  The map is a dict, as the reference's HashMap; the
    oracle's find_empty_neighbor and mutate are the only primitives.
    Returns (cells, rules, stats)."""
    cells = [c.copy() for c in cells]
    rules = [r.copy() for r in rules]
    cell_map = {tuple(int(v) for v in c["pos"]): i for i, c in enumerate(cells)}
    dense = np.full(dims[0] * dims[1] * dims[2], -1, np.int32)  # the same map as the primitive reads it
    for p, i in cell_map.items():
        dense[_vox(dims, p)] = i
    stream = (TAG_DIVISION << 32) | (tick & 0xFFFFFFFF)
    births, deaths = [], []
    tick_divisions = tick_deaths = 0
    for i, event in enumerate(events):  # :223-252
        if event == EVENT_DIVISION:
            target, draws = orc.find_empty_neighbor(cells[i]["pos"], dense, dims, sim, sim.rng_seed, stream, i)
            if target is not None:
                daughter = cells[i].copy()
                daughter["pos"] = target
                daughter["age"] = 0
                daughter["prep_remaining"] = 0
                rs = Stream(orc, sim.rng_seed, stream, i, draws)
                daughter["lineage_id"] = rs.u64()
                for k in range(16):
                    daughter["internal"][k] = F(daughter["internal"][k]) * F(0.5)
                    cells[i]["internal"][k] = F(cells[i]["internal"][k]) * F(0.5)
                drule, _ = orc.mutate(rules[i], sim, sim.rng_seed, stream, i, rs.draws)
                births.append((daughter, drule))
                tick_divisions += 1
        elif event == EVENT_DEATH:
            deaths.append(i)
            tick_deaths += 1
    deaths = sorted(set(deaths))  # :257-259
    for i in reversed(deaths):  # :260-267
        pos = tuple(int(v) for v in cells[i]["pos"])
        del cell_map[pos]
        cells[i] = cells[-1]
        rules[i] = rules[-1]
        cells.pop()
        rules.pop()
        if i < len(cells):
            cell_map[tuple(int(v) for v in cells[i]["pos"])] = i
    added = 0
    for c, r in births:  # :270-277
        pos = tuple(int(v) for v in c["pos"])
        if pos not in cell_map:
            cell_map[pos] = len(cells)
            cells.append(c)
            rules.append(r)
            added += 1
    stats = {"divisions": tick_divisions, "deaths": tick_deaths, "births_added": added, "n_cells": len(cells)}
    return cells, rules, stats


def seed_cells(orc, dims, seed, band, count, margin, z_lo, z_hi, cell_map):
    """seed_cells — crates/marl-sim/src/seeding.rs:28-58: up to count * 8
    attempts; x, y in [margin, dim - margin), z in [z_lo, min(z_hi, Z)); a taken
    voxel is skipped; each placed cell draws its u64 lineage.  cell_map (dict,
    updated) maps positions to indices; new indices continue from len(cell_map).
    Returns [(pos, lineage)]."""
    rs = Stream(orc, seed, (TAG_SEEDING << 32) | band, 0)
    out = []
    seeded = 0
    for _ in range(count * 8):
        if seeded >= count:
            break
        x = margin + rs.range(dims[0] - 2 * margin)
        y = margin + rs.range(dims[1] - 2 * margin)
        z = z_lo + rs.range(min(z_hi, dims[2]) - z_lo)
        pos = (x, y, z)
        if pos in cell_map:
            continue
        lineage = rs.u64()
        cell_map[pos] = len(cell_map)
        out.append((pos, lineage))
        seeded += 1
    return out
