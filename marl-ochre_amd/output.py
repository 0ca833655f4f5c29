"""Run outputs from the resident GPU state, in the reference's file formats.

The reductions and the record packing run on the GPU (`MarlGpu.cell_summary`,
`chemistry_profile`, `download_viewer_cells`, `sample`); this module only
formats and writes:

* ``ticks.csv`` — ≙ DataLogger::new / log_tick (crates/marl-output/src/data.rs:143-275);
* ``chem_<T>.csv`` — ≙ DataLogger::snapshot_chemistry (data.rs:292-335);
* the console lines of print_stats / print_z_profile (crates/marl-sim/src/stats.rs:9-103);
* ``tick_<T>.field.bin[.zst]``, ``tick_<T>.cells.bin[.zst]`` and ``run_meta.json`` —
  ≙ write_field_dump / write_cell_dump / write_run_meta
  (crates/marl-output/src/binary_dump.rs:66-120, 362-429; formats in
  crates/marl-format/src/lib.rs:25-93, 123-230, 350-364).

Averages come from FP64 sums computed in a different order from the
reference's sequential loops, so a printed value can differ only where the
rounding of the last printed digit is decided below FP64 precision.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from ._abi import S_EXT

ENDIANNESS = "little"
FIELD_DTYPE = "f32"
FIELD_LAYOUT = "z_y_x_species"
CELL_RECORD_STRIDE = 25
BINARY_FORMAT_VERSION = 2
CELL_RECORD_LAYOUT = "pos:f32[3],lineage_id:u64,starter_type:u8,energy:f32"
RULESET_LAYER_RECORD_LAYOUT = (
    "z:u16,reserved:u16,cell_count:u32,receptors:{k_half:f32,n_hill:f32,gain:f32}[s_receptors],"
    "transport:{uptake_rate:f32,secrete_rate:f32}[s_transporters],reactions:{k_m:f32,v_max:f32,k_cat:f32}[r_max],"
    "effectors:{threshold:f32,rate:f32}[s_effectors],fate:f32[4],hgt_propensity:f32,mutation_rate:f32")
VIEWER_DTYPE = np.dtype({"names": ["pos", "lineage_id", "starter_type", "energy"],
                         "formats": [(np.float32, 3), np.uint64, np.uint8, np.float32],
                         "offsets": [0, 12, 20, 21], "itemsize": 25})


U64_MAX = (1 << 64) - 1


class FormatError(ValueError):
    """A run_meta.json that does not describe this format (≙ FormatError, marl-format/src/lib.rs)."""


def field_byte_len(grid_x: int, grid_y: int, grid_z: int, s_ext: int):
    """Bytes of a raw field dump, None for a zero dimension or u64 overflow
    (marl-format/src/lib.rs:290-300)."""
    if min(grid_x, grid_y, grid_z, s_ext) <= 0:
        return None
    n = grid_x * grid_y * grid_z * s_ext * 4
    return n if n <= U64_MAX else None


def ruleset_layer_value_count(s_receptors=8, s_transporters=8, r_max=16, s_effectors=8) -> int:
    """f32 averages in one per-layer ruleset record (marl-format/src/lib.rs:303-318)."""
    return 3 * s_receptors + 2 * s_transporters + 3 * r_max + 2 * s_effectors + 6


def ruleset_layer_record_stride(s_receptors=8, s_transporters=8, r_max=16, s_effectors=8) -> int:
    return 8 + 4 * ruleset_layer_value_count(s_receptors, s_transporters, r_max, s_effectors)


def validate_run_meta(meta: dict) -> None:
    """The checks of RunMeta::validate, first mismatch raised
    (marl-format/src/lib.rs:220-285); keys a legacy file lacks take the
    defaults serde fills in (lib.rs:151-184)."""
    m = {"binary_format_version": 1, "binary_compression": "none", "field_file_pattern": "tick_<T>.field.bin",
         "cell_file_pattern": "tick_<T>.cells.bin", **meta}
    for key, want in (("endianness", ENDIANNESS), ("field_dtype", FIELD_DTYPE), ("field_layout", FIELD_LAYOUT),
                      ("cell_record_stride", CELL_RECORD_STRIDE)):
        if m[key] != want:
            raise FormatError(f"{key} mismatch: expected {want}, got {m[key]}")
    v = m["binary_format_version"]
    if v == 0 or v > BINARY_FORMAT_VERSION:
        raise FormatError(f"binary_format_version mismatch: expected 1..={BINARY_FORMAT_VERSION}, got {v}")
    if m["binary_compression"] not in ("none", "zstd"):
        raise FormatError(f"binary_compression mismatch: expected none or zstd, got {m['binary_compression']}")
    for key in ("field_file_pattern", "cell_file_pattern"):
        if not m[key]:
            raise FormatError(f"{key} must not be empty")
    want = field_byte_len(m["grid_x"], m["grid_y"], m["grid_z"], m["s_ext"])
    if want is not None and m["field_byte_len"] != want:
        raise FormatError(f"field_byte_len mismatch: expected {want}, got {m['field_byte_len']}")


def read_snapshot(run_dir, tick: int):
    """Read a snapshot back as the viewer does (marl-viewer-core/src/io.rs):
    run_meta.json, the field as float32 [Z, Y, X, S_EXT] and the cell records
    (VIEWER_DTYPE), after the size checks of scripts/check_binary_dump.py."""
    run_dir = Path(run_dir)
    meta = json.loads((run_dir / "run_meta.json").read_text(encoding="utf-8"))
    validate_run_meta(meta)
    fpat = meta.get("field_file_pattern", "tick_<T>.field.bin").replace("<T>", str(tick))
    cpat = meta.get("cell_file_pattern", "tick_<T>.cells.bin").replace("<T>", str(tick))
    fb = read_payload(run_dir / fpat)
    if len(fb) != meta["field_byte_len"]:
        raise FormatError(f"field size mismatch: got {len(fb)}, expected {meta['field_byte_len']}")
    cb = read_payload(run_dir / cpat)
    stride = meta["cell_record_stride"]
    if len(cb) % stride:
        raise FormatError(f"cell file size {len(cb)} is not divisible by stride {stride}")
    field = np.frombuffer(fb, "<f4").reshape(meta["grid_z"], meta["grid_y"], meta["grid_x"], meta["s_ext"])
    return meta, field, np.frombuffer(cb, VIEWER_DTYPE)


def _zstd():
    import pyarrow as pa

    if not pa.Codec.is_available("zstd"):
        raise RuntimeError("zstd compression is not available (pyarrow without zstd)")
    return pa


def _write_payload(path: Path, payload: bytes, compression: str, level: int) -> None:
    """Write via a .tmp file and rename, like write_binary_payload (binary_dump.rs:80-107)."""
    tmp = Path(str(path) + ".tmp")
    try:
        if compression == "zstd":
            pa = _zstd()
            payload = pa.Codec("zstd", compression_level=level or None).compress(payload, asbytes=True)
        elif compression != "none":
            raise ValueError(f"unknown compression {compression!r}")
        tmp.write_bytes(payload)
        os.replace(tmp, path)
    except BaseException:
        if tmp.exists():
            tmp.unlink()
        raise


def read_payload(path: Path) -> bytes:
    """Read a dump written here (decompressing .zst)."""
    data = Path(path).read_bytes()
    if str(path).endswith(".zst"):
        pa = _zstd()
        with pa.CompressedInputStream(pa.BufferReader(data), "zstd") as f:
            return f.read()
    return data


def binary_file_path(out_dir, tick: int, stem: str, compression: str = "none") -> Path:
    suffix = f"tick_{tick}.{stem}.bin" + (".zst" if compression == "zstd" else "")
    return Path(out_dir) / suffix


def write_field_dump(gpu, tick: int, out_dir, compression: str = "none", level: int = 3) -> Path:
    """The AoS f32 field ("z_y_x_species"), little-endian (binary_dump.rs:109-113)."""
    Path(out_dir).mkdir(parents=True, exist_ok=True)
    path = binary_file_path(out_dir, tick, "field", compression)
    _write_payload(path, gpu.download_field_aos().astype("<f4", copy=False).tobytes(), compression, level)
    return path


def write_cell_dump(gpu, tick: int, out_dir, compression: str = "none", level: int = 3) -> Path:
    """25-byte viewer records packed on the GPU (binary_dump.rs:115-120)."""
    Path(out_dir).mkdir(parents=True, exist_ok=True)
    path = binary_file_path(out_dir, tick, "cells", compression)
    _write_payload(path, gpu.download_viewer_cells(), compression, level)
    return path


def run_meta(dims, write_binary_field=True, write_binary_cells=True, compression="none", level=3,
             snapshot_interval=500, max_ticks=5000, ruleset_interval=0, ruleset_output_mode="off") -> dict:
    """The run_meta.json content of write_run_meta (binary_dump.rs:362-429)
    without the full-ruleset keys (the full ruleset dump is not written here)."""
    X, Y, Z = dims
    ext = ".zst" if compression == "zstd" else ""
    field_count = X * Y * Z * S_EXT
    layer_values = 3 * 8 + 2 * 8 + 3 * 16 + 2 * 8 + 6
    return {
        "grid_x": X, "grid_y": Y, "grid_z": Z, "s_ext": S_EXT, "m_int": 16,
        "field_dtype": FIELD_DTYPE, "field_layout": FIELD_LAYOUT, "field_byte_len": field_count * 4,
        "cell_record_stride": CELL_RECORD_STRIDE, "endianness": ENDIANNESS,
        "binary_format_version": BINARY_FORMAT_VERSION, "binary_compression": compression,
        "binary_compression_level": level if compression == "zstd" else 0,
        "write_binary_field": bool(write_binary_field), "write_binary_cells": bool(write_binary_cells),
        "field_file_pattern": "tick_<T>.field.bin" + ext, "cell_file_pattern": "tick_<T>.cells.bin" + ext,
        "snapshot_interval": snapshot_interval, "max_ticks": max_ticks, "field_count": field_count,
        "cell_header_byte_len": 0, "cell_count_source": "file_size_divided_by_cell_record_stride",
        "cell_pos_units": "grid_voxel_indices", "cell_record_layout": CELL_RECORD_LAYOUT,
        "ruleset_interval": ruleset_interval, "ruleset_output_mode": ruleset_output_mode,
        "ruleset_layer_file_pattern": "tick_<T>.ruleset_layers.bin" + ext,
        "ruleset_layer_record_stride": 8 + 4 * layer_values,
        "ruleset_layer_record_layout": RULESET_LAYER_RECORD_LAYOUT,
    }


def write_run_meta(out_dir, dims, **kw) -> Path:
    Path(out_dir).mkdir(parents=True, exist_ok=True)
    path = Path(out_dir) / "run_meta.json"
    path.write_text(json.dumps(run_meta(dims, **kw), indent=2))
    return path


class TickLog:
    """ticks.csv (≙ DataLogger with write_tick_log, data.rs:143-275)."""

    def __init__(self, out_dir, grid_z: int):
        Path(out_dir).mkdir(parents=True, exist_ok=True)
        self.path = Path(out_dir) / "ticks.csv"
        self.f = open(self.path, "w")
        self.f.write("tick,population,avg_energy,avg_enzyme_a,avg_enzyme_b,avg_active_rxns,divisions_this_tick,"
                     "deaths_this_tick" + "".join(f",z{z}_cells" for z in range(grid_z)) + "\n")
        self.f.flush()

    @staticmethod
    def row(tick: int, summary: dict, z_counts, divisions: int, deaths: int) -> str:
        pop = summary["n_cells"]
        if pop > 0:
            e, a, b = (summary[k] / pop for k in ("sum_energy", "sum_enzyme_a", "sum_enzyme_b"))
            r = summary["sum_active_reactions"] / pop
        else:
            e = a = b = r = 0.0
        return (f"{tick},{pop},{e:.6f},{a:.6f},{b:.6f},{r:.4f},{divisions},{deaths}"
                + "".join(f",{int(c)}" for c in z_counts) + "\n")

    def log_tick(self, gpu, tick: int, divisions: int, deaths: int) -> None:
        summary, zc = gpu.cell_summary()
        self.f.write(self.row(tick, summary, zc, divisions, deaths))
        self.f.flush()

    def close(self) -> None:
        self.f.close()


def chemistry_csv(profile: np.ndarray) -> str:
    lines = ["z" + "".join(f",species_{s}_avg" for s in range(S_EXT)) + ",light_avg"]
    for z, row in enumerate(profile):
        lines.append(f"{z}" + "".join(f",{v:.6f}" for v in row))
    return "\n".join(lines) + "\n"


def snapshot_chemistry(gpu, tick: int, out_dir) -> Path:
    path = Path(out_dir) / f"chem_{tick}.csv"
    path.write_text(chemistry_csv(gpu.chemistry_profile()))
    return path


def stats_line(gpu, tick: int, divisions: int, deaths: int, elapsed_s: float) -> str:
    """The line print_stats prints (stats.rs:9-82).

    The averages here divide the GPU's FP64 sums; print_stats sums FP32 values
    one after another, so on large populations its last printed digit can
    differ by the FP32 accumulation error it carries."""
    X, Y, Z = gpu.dims
    summary, _ = gpu.cell_summary()
    n = summary["n_cells"]
    if n == 0:
        return f"t={tick:>5} | EXTINCT | div={divisions} death={deaths}"
    cx, cy = X // 2, Y // 2
    ox_top, ox_mid, red_bot, org_mid = gpu.sample([(cx, cy, 0, 1), (cx, cy, Z // 2, 1), (cx, cy, Z - 1, 2),
                                                   (cx, cy, Z // 2, 4)])
    top, mid, bot = summary["z_thirds"]
    tps = (tick + 1) / elapsed_s if elapsed_s > 0 else 0.0
    return (f"t={tick:>5} | pop={n:>6} (top:{top:>5} mid:{mid:>5} bot:{bot:>5}) | "
            f"E={summary['sum_energy'] / n:.2f} enz={summary['sum_enzyme_a'] / n:.3f} "
            f"rxn={summary['sum_active_reactions'] / n:.1f} | ox={ox_top:.2f}/{ox_mid:.2f} red={red_bot:.2f} "
            f"org={org_mid:.2f} | {tps:.1f} t/s")


def z_profile_text(gpu) -> str:
    """The table print_z_profile prints (stats.rs:84-103)."""
    X, Y, Z = gpu.dims
    cx, cy = X // 2, Y // 2
    _, zc = gpu.cell_summary()
    q = []
    for z in range(Z):
        q += [(cx, cy, z, -1), (cx, cy, z, 1), (cx, cy, z, 2), (cx, cy, z, 3), (cx, cy, z, 4)]
    v = gpu.sample(q).reshape(Z, 5)
    lines = [f"{'z':>3} {'cells':>6} {'light':>6} {'oxidnt':>6} {'reduct':>6} {'carbon':>6} {'organic':>6}"]
    for z in range(Z):
        l, ox, re, ca, og = v[z]
        lines.append(f"{z:>3} {int(zc[z]):>6} {l:>6.3f} {ox:>6.3f} {re:>6.3f} {ca:>6.3f} {og:>6.3f}")
    return "\n".join(lines) + "\n"
