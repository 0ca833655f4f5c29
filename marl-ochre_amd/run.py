"""Command-line run of the simulation on the GPU.

Mirrors the reference's `marl-engine` binary (crates/marl-engine/src/main.rs:1-8):
configuration by `Config::load` (crates/marl-config/src/lib.rs:300-376: built-in
defaults, then a TOML file, `marl.toml` or `--config <path>`, then run-control
flags), and the tick loop with its outputs by `marl_sim::run`
(crates/marl-sim/src/lib.rs:34-420): field priming, three seeded metabolisms,
per-tick logging, periodic stats, binary and CSV snapshots, ruleset dumps, the
final z-profile and `summary.md`.

Differences, by design:
- Every tick runs on the GPU (`marl_tick_graph`, or `marl_tick` with `--eager`).
  `--gpu-diffusion` is accepted and changes nothing. The reference falls back to
  its CPU diffusion when the GPU is missing or fails (lib.rs:133-146, 180-191);
  this program has no CPU path, so it reports the error and exits with status 1.
- The grid is a runtime choice (`--grid XxYxZ` or a `[grid]` table; the
  reference fixes 128x128x64 at compile time) and the RNG is Philox keyed by
  `--rng-seed` (the reference draws from the OS, so its runs do not repeat).
- PPM image snapshots (crates/marl-output/src/snapshot.rs) are not written;
  a configuration asking for them gets one warning.
"""
from __future__ import annotations

import dataclasses
import os
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import output as out_mod
from ._abi import CELL_DTYPE, RULESET_DTYPE, S_EXT, GpuError
from .config import SimulationConfig

COMPRESSIONS = ("none", "zstd")
RULESET_MODES = ("off", "layer_averages", "full", "both")


@dataclass
class OutputConfig:
    """≙ OutputConfig and its defaults (crates/marl-config/src/lib.rs:230-280)."""
    max_ticks: int = 5000
    stats_interval: int = 100
    snapshot_interval: int = 500
    image_interval: int = 500
    seed_count: int = 30
    output_dir: str = ""  # default output/run_<X>x<Y>x<Z>, filled in from the grid
    xz_snapshot_species: list = field(default_factory=list)
    xy_slice_depths_frac: list = field(default_factory=list)
    write_binary_field: bool = True
    write_binary_cells: bool = True
    binary_compression: str = "zstd"
    binary_compression_level: int = 3
    ruleset_interval: int = 1000
    ruleset_output_mode: str = "off"
    write_tick_log: bool = False
    write_csv_snapshots: bool = False
    write_ancestry_map: bool = False
    write_density_map: bool = False

    def writes_images(self) -> bool:  # lib.rs:151-154
        return bool(self.xz_snapshot_species or self.xy_slice_depths_frac or self.write_density_map
                    or self.write_ancestry_map)


@dataclass
class Config:
    simulation: SimulationConfig = field(default_factory=SimulationConfig)
    output: OutputConfig = field(default_factory=OutputConfig)
    grid: tuple = (128, 128, 64)  # the reference's GRID_X, GRID_Y, GRID_Z (crates/marl-config/src/lib.rs:17-20)
    rng_seed: int | None = None
    eager: bool = False


class ConfigError(ValueError):
    pass


# Field kinds of the TOML tables, as serde reads them into the Rust types.
_U = "uint"
_F = "float"
_B = "bool"
_S = "str"
_SIM_KINDS = {f.name: (_U if f.type in ("int",) else _F) for f in dataclasses.fields(SimulationConfig)}
_SIM_KINDS.update({"d_voxel": ("floats", S_EXT), "lambda_decay": ("floats", S_EXT),
                   "seed_margin": ("uint_max", 0xFFFF), "division_neighbor_distance": ("uint_max", 0xFF)})
_SIM_KINDS.pop("rng_seed")
_OUT_KINDS = {
    "max_ticks": _U, "stats_interval": _U, "snapshot_interval": _U, "image_interval": _U, "seed_count": _U,
    "output_dir": _S, "xz_snapshot_species": ("uints", None), "xy_slice_depths_frac": ("floats", None),
    "write_binary_field": _B, "write_binary_cells": _B, "binary_compression": ("enum", COMPRESSIONS),
    "binary_compression_level": "int", "ruleset_interval": _U, "ruleset_output_mode": ("enum", RULESET_MODES),
    "write_tick_log": _B, "write_csv_snapshots": _B, "write_ancestry_map": _B, "write_density_map": _B,
}


def _convert(table: str, key: str, kind, v):
    where = f"{table}.{key}"
    if kind == _F:
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise ConfigError(f"invalid type for {where}: expected a number")
        return float(v)  # serde accepts TOML integers for f32 fields
    if kind in (_U, "int"):
        if isinstance(v, bool) or not isinstance(v, int):
            raise ConfigError(f"invalid type for {where}: expected an integer")
        if kind == _U and v < 0:
            raise ConfigError(f"invalid value for {where}: expected a non-negative integer")
        return v
    if kind == _B:
        if not isinstance(v, bool):
            raise ConfigError(f"invalid type for {where}: expected a boolean")
        return v
    if kind == _S:
        if not isinstance(v, str):
            raise ConfigError(f"invalid type for {where}: expected a string")
        return v
    name, arg = kind
    if name == "uint_max":  # u16 / u8 fields
        v = _convert(table, key, _U, v)
        if v > arg:
            raise ConfigError(f"invalid value for {where}: expected at most {arg}")
        return v
    if name == "enum":
        if v not in arg:
            raise ConfigError(f"unknown variant {v!r} for {where}, expected one of {', '.join(arg)}")
        return v
    if not isinstance(v, list):
        raise ConfigError(f"invalid type for {where}: expected an array")
    if arg is not None and len(v) != arg:
        raise ConfigError(f"invalid length {len(v)} for {where}, expected {arg}")
    return [_convert(table, key, _F if name == "floats" else _U, x) for x in v]


def parse_toml(text: str) -> Config:
    """A TOML document into a Config: tables `[simulation]` and `[output]` with
    the reference's field names; missing keys keep their defaults, unknown keys
    are ignored (serde without deny_unknown_fields), a value of the wrong type
    rejects the whole file.  `[grid]` (x, y, z) is this project's addition."""
    try:
        import tomllib as toml  # Python >= 3.11
    except ModuleNotFoundError:
        import tomli as toml
    try:
        doc = toml.loads(text)
    except toml.TOMLDecodeError as e:
        raise ConfigError(str(e)) from e
    cfg = Config()
    for table, kinds, target in (("simulation", _SIM_KINDS, cfg.simulation), ("output", _OUT_KINDS, cfg.output)):
        t = doc.get(table, {})
        if not isinstance(t, dict):
            raise ConfigError(f"invalid type for {table}: expected a table")
        for k, v in t.items():
            if k in kinds:
                setattr(target, k, _convert(table, k, kinds[k], v))
    g = doc.get("grid")
    if g is not None:
        if not isinstance(g, dict):
            raise ConfigError("invalid type for grid: expected a table")
        cfg.grid = tuple(_convert("grid", k, _U, g.get(k, d)) for k, d in zip("xyz", cfg.grid))
    return cfg


def _parse_grid(s: str):
    parts = s.lower().split("x")
    if len(parts) != 3 or not all(p.isdigit() for p in parts):
        return None
    return tuple(int(p) for p in parts)


def load(argv, cwd=".", warn=None) -> Config:
    """≙ Config::load (crates/marl-config/src/lib.rs:300-376).

    1. built-in defaults;
    2. the TOML file named by `--config <path>`, else `marl.toml` in cwd, when
       it can be read; a file that fails to parse prints a warning and leaves the
       defaults in place;
    3. run-control flags: --ticks, --stats, --snapshot, --ruleset-interval,
       --images, --seed, --output.  A flag without a value, or whose value does
       not parse, is skipped.

    Added here: --grid XxYxZ, --rng-seed N, --eager, --gpu-diffusion (no-op).
    """
    warn = warn or (lambda m: print(m, file=sys.stderr))
    args = list(argv)
    cfg = Config()
    config_path = None
    i = 0
    while i < len(args):
        if args[i] == "--config" and i + 1 < len(args):
            config_path = args[i + 1]
            i += 2
        else:
            i += 1
    path = Path(config_path) if config_path is not None else Path(cwd) / "marl.toml"
    if config_path is not None and not Path(config_path).is_absolute():
        path = Path(cwd) / config_path
    try:
        text = path.read_text()
    except OSError:
        text = None
    if text is not None:
        try:
            cfg = parse_toml(text)
        except ConfigError as e:
            warn(f"Warning: failed to parse config {config_path or 'marl.toml'}: {e}")

    def as_uint(v):  # Rust's integer parse: digits with an optional leading '+'
        d = v[1:] if v.startswith("+") else v
        return int(d) if d.isdigit() else None

    flags = {"--ticks": "max_ticks", "--stats": "stats_interval", "--snapshot": "snapshot_interval",
             "--ruleset-interval": "ruleset_interval", "--images": "image_interval", "--seed": "seed_count"}
    i = 0
    while i < len(args):
        a = args[i]
        if a == "--config":
            i += 2
        elif a in flags and i + 1 < len(args):
            v = as_uint(args[i + 1])
            if v is not None:
                setattr(cfg.output, flags[a], v)
            i += 2
        elif a == "--output" and i + 1 < len(args):
            cfg.output.output_dir = args[i + 1]
            i += 2
        elif a == "--grid" and i + 1 < len(args):
            g = _parse_grid(args[i + 1])
            if g is not None:
                cfg.grid = g
            i += 2
        elif a == "--rng-seed" and i + 1 < len(args):
            v = as_uint(args[i + 1])
            if v is not None:
                cfg.rng_seed = v
            i += 2
        elif a == "--eager":
            cfg.eager = True
            i += 1
        else:
            i += 1
    if not cfg.output.output_dir:
        cfg.output.output_dir = "output/run_{}x{}x{}".format(*cfg.grid)
    return cfg


# ---------------------------------------------------------------------------
# Setup: field priming and seeding (crates/marl-sim/src/seeding.rs:9-58,
# crates/marl-sim/src/lib.rs:57-105).

def primed_field(dims, sim: SimulationConfig) -> np.ndarray:
    """≙ init_field_boundaries (seeding.rs:9-26): an empty AoS field with the top
    `boundary_prime_layers` holding oxidant and carbon and the bottom ones
    reductant."""
    X, Y, Z = dims
    f = np.zeros((Z, Y, X, S_EXT), np.float32)
    layers = sim.boundary_prime_layers
    f[:layers, :, :, 1] = np.float32(sim.boundary_prime_oxidant)
    f[:layers, :, :, 3] = np.float32(sim.boundary_prime_carbon)
    f[max(Z - layers, 0):, :, :, 2] = np.float32(sim.boundary_prime_reductant)
    return f.reshape(-1)


def seed_bands(dims, sim: SimulationConfig):
    """The three depth bands of run() (lib.rs:62-105), computed as the reference
    does: canonical 200-layer depths scaled by Z/200 in f32, truncated to u16,
    with the upper bound widened to at least lo+1 / lo+3 / lo+3."""
    f32 = np.float32
    z_scale = f32(dims[2]) / f32(200.0)

    def u16(x):
        return int(min(max(float(x), 0.0), 65535.0))

    out = []
    for lo_f, hi_f, widen in ((sim.phototroph_z_lo, sim.phototroph_z_hi, 1.0),
                              (sim.chemolithotroph_z_lo, sim.chemolithotroph_z_hi, 3.0),
                              (sim.anaerobe_z_lo, sim.anaerobe_z_hi, 3.0)):
        lo = u16(f32(lo_f) * z_scale)
        hi = u16(max(f32(hi_f) * z_scale, f32(lo) + f32(widen)))
        out.append((lo, hi))
    return out


def seed_population(dims, sim: SimulationConfig, seed_count: int, rng_seed: int):
    """seed_cells for phototrophs, chemolithotrophs and anaerobes, in that order,
    each with its starter's ruleset and initial state (starter_metabolisms.rs)."""
    from .scenarios import seed_positions
    from .starters import make_cell, make_ruleset

    nv = dims[0] * dims[1] * dims[2]
    ca = np.full(nv, -1, np.int32)
    cells, rules = [], []
    for kind, (lo, hi) in enumerate(seed_bands(dims, sim)):
        base = sum(len(c) for c in cells)
        pos, lin, ca = seed_positions(dims, rng_seed, kind, seed_count, sim.seed_margin, lo, hi, ca,
                                      index_base=base)
        c = np.zeros(len(pos), CELL_DTYPE)
        r = np.zeros(len(pos), RULESET_DTYPE)
        r[:] = make_ruleset(kind)
        for k in range(len(pos)):
            c[k] = make_cell(kind, pos[k], int(lin[k]))
        cells.append(c)
        rules.append(r)
    return np.concatenate(cells), np.concatenate(rules)


# ---------------------------------------------------------------------------

def run(cfg: Config, stdout=None) -> int:
    """≙ marl_sim::run (lib.rs:34-420) on the GPU.  Returns the process status."""
    from .engine import MarlGpu

    p = (lambda *a: print(*a, file=stdout, flush=True)) if stdout else (lambda *a: print(*a, flush=True))
    warn = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    o = cfg.output
    dims = tuple(cfg.grid)
    seed = cfg.rng_seed if cfg.rng_seed is not None else int.from_bytes(os.urandom(8), "little") >> 1
    sim = cfg.simulation.replace(rng_seed=seed)
    Path(o.output_dir).mkdir(parents=True, exist_ok=True)
    log = out_mod.TickLog(o.output_dir, dims[2]) if o.write_tick_log else None
    mode = o.ruleset_output_mode
    if o.write_binary_field or o.write_binary_cells or mode != "off":
        out_mod.write_run_meta(o.output_dir, dims, write_binary_field=o.write_binary_field,
                               write_binary_cells=o.write_binary_cells, compression=o.binary_compression,
                               level=o.binary_compression_level, snapshot_interval=o.snapshot_interval,
                               max_ticks=o.max_ticks, ruleset_interval=o.ruleset_interval,
                               ruleset_output_mode=mode)
    field0 = primed_field(dims, sim)
    cells, rules = seed_population(dims, sim, o.seed_count, seed)

    X, Y, Z = dims
    p("MARL v0.3 — MI355X (gfx950) path (Winogradsky)")
    p(f"Grid: {X}x{Y}x{Z} ({X * Y * Z / 1e6:.1f}M voxels), Species: {S_EXT} ext / 16 int")
    p(f"Seeded {len(cells)} cells (photo/chemo/anaerobe)")
    p(f"Output: {o.output_dir}")
    p(f"Plan: {o.max_ticks} ticks, stats every {o.stats_interval}, snapshots every {o.snapshot_interval}, "
      f"images every {o.image_interval}")
    p(f"Diffusion: GPU (rng seed {seed})")
    p("---")
    if o.writes_images():
        warn("Warning: image snapshots are not produced by this program; skipping them")
    try:
        g = MarlGpu(dims)
    except GpuError as e:
        warn(f"Error: GPU unavailable ({e}); this program has no CPU path")
        return 1
    total_div = total_death = 0
    registry = out_mod.ReactionRegistry()
    start = time.perf_counter()
    last = o.max_ticks - 1
    try:
        g.upload_field_aos(field0)
        g.upload_cells(cells, rules)
        step = g.tick if cfg.eager else g.tick_graph
        for tick in range(o.max_ticks):
            st = step(sim, tick, 1)
            div, death = int(st["divisions"]), int(st["deaths"])
            total_div += div
            total_death += death
            if log is not None:
                log.log_tick(g, tick, div, death)
            if tick % max(o.stats_interval, 1) == 0 or tick == last:
                p(out_mod.stats_line(g, tick, total_div, total_death, time.perf_counter() - start))
            if tick % max(o.snapshot_interval, 1) == 0 or tick == last:
                _snapshot(g, tick, o, registry, warn)
            if mode != "off" and (tick % max(o.ruleset_interval, 1) == 0 or tick == last):
                _ruleset_dumps(g, tick, o, warn)
        p("\n=== FINAL Z-LAYER PROFILE ===")
        p(out_mod.z_profile_text(g).rstrip("\n"))
        runtime = time.perf_counter() - start
        n = g.cell_count()
        p(f"\nDone. {o.max_ticks} ticks in {runtime:.1f}s, final pop={n}, div={total_div}, death={total_death}")
        try:
            out_mod.write_summary(g, o.output_dir, o.max_ticks, runtime, total_div, total_death, sim)
            p(f"Summary written to {o.output_dir}/summary.md")
        except OSError as e:
            warn(f"Warning: failed to write summary: {e}")
        if o.write_csv_snapshots:
            try:
                registry.write(o.output_dir)
                p(f"Reaction registry written to {o.output_dir}/reaction_registry.csv ({registry.count()} topologies)")
            except OSError as e:
                warn(f"Warning: failed to write reaction registry: {e}")
    except GpuError as e:
        warn(f"Error: GPU failure ({e}); this program has no CPU path")
        return 1
    finally:
        if log is not None:
            log.close()
        g.close()
    return 0


def _snapshot(g, tick, o: OutputConfig, registry, warn) -> None:
    """Binary and optional CSV snapshots (lib.rs:300-345)."""
    comp, lvl = o.binary_compression, o.binary_compression_level
    if o.write_binary_field:
        try:
            out_mod.write_field_dump(g, tick, o.output_dir, comp, lvl)
        except OSError as e:
            warn(f"Warning: failed to write binary field snapshot at tick {tick}: {e}")
    if o.write_binary_cells:
        try:
            out_mod.write_cell_dump(g, tick, o.output_dir, comp, lvl)
        except OSError as e:
            warn(f"Warning: failed to write binary cell snapshot at tick {tick}: {e}")
    if o.write_csv_snapshots:
        cells, rules = g.download_cells()
        for what, fn in (("chemistry", lambda: out_mod.snapshot_chemistry(g, tick, o.output_dir)),
                         ("cell", lambda: out_mod.snapshot_cells(cells, rules, tick, o.output_dir)),
                         ("reaction", lambda: out_mod.snapshot_reactions(cells, rules, tick, o.output_dir, registry))):
            try:
                fn()
            except OSError as e:
                warn(f"Warning: failed to write {what} snapshot at tick {tick}: {e}")


def _ruleset_dumps(g, tick, o: OutputConfig, warn) -> None:
    """Ruleset layer averages and the deduplicated full dump (lib.rs:346-370)."""
    cells, rules = g.download_cells()
    cells_pos = cells["pos"]
    mode = o.ruleset_output_mode
    if mode in ("layer_averages", "both"):
        try:
            out_mod.write_ruleset_layer_dump(rules, cells_pos, g.dims[2], tick, o.output_dir,
                                             o.binary_compression, o.binary_compression_level)
        except OSError as e:
            warn(f"Warning: failed to write binary ruleset layer snapshot at tick {tick}: {e}")
    if mode in ("full", "both"):
        try:
            out_mod.write_ruleset_full_dump(rules, cells_pos, tick, o.output_dir, o.binary_compression,
                                            o.binary_compression_level)
        except OSError as e:
            warn(f"Warning: failed to write binary ruleset full dump at tick {tick}: {e}")


USAGE = """usage: python marl-ochre_amd [--config PATH] [--ticks N] [--stats N] [--snapshot N]
                          [--ruleset-interval N] [--images N] [--seed N] [--output DIR]
                          [--grid XxYxZ] [--rng-seed N] [--eager] [--gpu-diffusion]

Runs the simulation on the GPU with the reference engine's flags and marl.toml (defaults, then the
TOML file, then the flags; a flag without a parsable value is skipped). --help prints this text."""


def main(argv=None) -> int:
    args = sys.argv[1:] if argv is None else list(argv)
    if "--help" in args or "-h" in args:  # the reference has no help flag; this one only prints
        print(USAGE)
        return 0
    cfg = load(args)
    return run(cfg)
