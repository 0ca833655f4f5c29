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


# ---------------------------------------------------------------------------
# Cell and reaction CSV snapshots (crates/marl-output/src/data.rs:18-130, 347-466)

def rust_f32(v) -> str:
    """Rust's `Display` of an f32: the shortest decimal that reads back as the
    same f32, never in exponent form."""
    return np.format_float_positional(np.float32(v), unique=True, trim="-")


def active_reactions(rules) -> np.ndarray:
    """Reactions with |v_max| > 1e-9 per ruleset (the f32 comparison of data.rs:219-224)."""
    return (np.abs(rules["reactions"]["v_max"]) > np.float32(1e-9)).sum(axis=1)


class ReactionRegistry:
    """Reaction topology (substrate, product, catalyst, cofactor) -> id in
    first-seen order (≙ ReactionRegistry, data.rs:40-130)."""

    def __init__(self):
        self.ids = {}

    def register(self, topo) -> int:
        return self.ids.setdefault(tuple(int(t) for t in topo), len(self.ids))

    def count(self) -> int:
        return len(self.ids)

    def write(self, out_dir) -> Path:
        path = Path(out_dir) / "reaction_registry.csv"
        lines = ["reaction_id,substrate,product,catalyst,cofactor"]
        lines += [f"{i},{a},{b},{c},{d}" for (a, b, c, d), i in sorted(self.ids.items(), key=lambda kv: kv[1])]
        path.write_text("\n".join(lines) + "\n")
        return path


def cells_csv(cells, rules) -> str:
    """cells_<T>.csv rows (data.rs:347-395): internals with {:.6}."""
    act = active_reactions(rules)
    lines = ["x,y,z,energy,enzyme_a,enzyme_b,active_rxns,lineage_id,age,quiescent,starter_type"]
    for c, a in zip(cells, act):
        i = c["internal"]
        lines.append(f"{c['pos'][0]},{c['pos'][1]},{c['pos'][2]},{float(i[0]):.6f},{float(i[5]):.6f},"
                     f"{float(i[6]):.6f},{int(a)},{int(c['lineage_id'])},{int(c['age'])},{int(c['quiescent'] != 0)},"
                     f"{int(c['starter_type'])}")
    return "\n".join(lines) + "\n"


def snapshot_cells(cells, rules, tick: int, out_dir) -> Path:
    path = Path(out_dir) / f"cells_{tick}.csv"
    path.write_text(cells_csv(cells, rules))
    return path


def reactions_csv(cells, rules, registry: ReactionRegistry) -> str:
    """reactions_<T>.csv rows (data.rs:399-433): active reactions' registry ids,
    registered in cell order then slot order."""
    rx = rules["reactions"]
    lines = ["x,y,z,starter_type,lineage_id,reaction_ids"]
    for k, c in enumerate(cells):
        ids = [registry.register((r["substrate"], r["product"], r["catalyst"], r["cofactor"]))
               for r in rx[k] if abs(np.float32(r["v_max"])) > np.float32(1e-9)]
        lines.append(f"{c['pos'][0]},{c['pos'][1]},{c['pos'][2]},{int(c['starter_type'])},{int(c['lineage_id'])},"
                     + ";".join(str(i) for i in ids))
    return "\n".join(lines) + "\n"


def snapshot_reactions(cells, rules, tick: int, out_dir, registry: ReactionRegistry) -> Path:
    path = Path(out_dir) / f"reactions_{tick}.csv"
    path.write_text(reactions_csv(cells, rules, registry))
    return path


# ---------------------------------------------------------------------------
# Ruleset dumps (crates/marl-output/src/binary_dump.rs:122-328; record layouts
# crates/marl-format/src/lib.rs:52-93)

RULESET_FULL_MAGIC = b"MRSF"
RULESET_FULL_FORMAT_VERSION = 1
RULESET_FULL_HEADER_SIZE = 24
RULESET_FULL_CELL_REF_STRIDE = 10
RULESET_FULL_CANONICAL_SIZE = 536


def _canonical_dtype():
    names, formats, offsets = [], [], []
    off = 0

    def add(name, fmt, size):
        nonlocal off
        names.append(name); formats.append(fmt); offsets.append(off)
        off += size
    for k in range(8):
        add(f"rc{k}", ("<f4", 3), 12)
    for k in range(8):
        add(f"tr{k}", ("<f4", 2), 8); add(f"tr{k}_sp", ("u1", 2), 2)
    for k in range(16):
        add(f"rx{k}_w", ("u1", 4), 4); add(f"rx{k}", ("<f4", 3), 12)
    for k in range(8):
        add(f"ef{k}", ("<f4", 2), 8); add(f"ef{k}_sp", ("u1", 2), 2)
    add("fate", ("<f4", 4), 16)
    add("tail", ("<f4", 2), 8)
    assert off == RULESET_FULL_CANONICAL_SIZE
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": off})


CANONICAL_DTYPE = _canonical_dtype()


def canonical_rulesets(rules) -> np.ndarray:
    """Each ruleset's 536-byte canonical record (ruleset_to_canonical_bytes,
    binary_dump.rs:188-240): receptors {k_half, n_hill, gain}, transport
    {uptake, secrete, ext, int}, reactions {substrate, product, catalyst,
    cofactor, k_m, v_max, k_cat}, effectors {threshold, rate, int, ext}, fate,
    hgt_propensity, mutation_rate — little-endian, packed."""
    n = len(rules)
    r = np.zeros(n, CANONICAL_DTYPE)
    for k in range(8):
        rc = rules["receptors"][:, k]
        r[f"rc{k}"] = np.stack([rc["k_half"], rc["n_hill"], rc["gain"]], 1)
        t = rules["transport"][:, k]
        r[f"tr{k}"] = np.stack([t["uptake_rate"], t["secrete_rate"]], 1)
        r[f"tr{k}_sp"] = np.stack([t["ext_species"], t["int_species"]], 1)
        e = rules["effectors"][:, k]
        r[f"ef{k}"] = np.stack([e["threshold"], e["rate"]], 1)
        r[f"ef{k}_sp"] = np.stack([e["int_species"], e["ext_species"]], 1)
    for k in range(16):
        x = rules["reactions"][:, k]
        r[f"rx{k}_w"] = np.stack([x["substrate"], x["product"], x["catalyst"], x["cofactor"]], 1)
        r[f"rx{k}"] = np.stack([x["k_m"], x["v_max"], x["k_cat"]], 1)
    r["fate"] = np.stack([rules["division_energy"], rules["death_energy"], rules["quiescence_energy"],
                          rules["division_prep_ticks"]], 1)
    r["tail"] = np.stack([rules["hgt_propensity"], rules["mutation_rate"]], 1)
    return r


def _layer_values(records) -> np.ndarray:
    """The 110 averaged f32 values of a ruleset, in append_ruleset_layer_values
    order (binary_dump.rs:122-152), as an [n, 110] array."""
    cols = [records[f"rc{k}"] for k in range(8)] + [records[f"tr{k}"] for k in range(8)] + \
        [records[f"rx{k}"] for k in range(16)] + [records[f"ef{k}"] for k in range(8)] + \
        [records["fate"], records["tail"]]
    return np.concatenate(cols, axis=1)


def ruleset_layer_payload(records, z_of_cell, grid_z: int) -> bytes:
    """Per z: z:u16, 0:u16, count:u32, then the 110 layer means as f32 (FP64
    sums in cell order, divided, rounded to f32; zeros for an empty layer)."""
    vals = _layer_values(records).astype(np.float64)
    z = np.asarray(z_of_cell, np.int64)
    counts = np.bincount(z, minlength=grid_z)[:grid_z]
    out = bytearray()
    sums = np.stack([np.bincount(z, weights=vals[:, k], minlength=grid_z)[:grid_z] for k in range(vals.shape[1])], 1) \
        if len(z) else np.zeros((grid_z, vals.shape[1]))
    for zz in range(grid_z):
        c = int(counts[zz])
        out += np.array([zz, 0], "<u2").tobytes() + np.array([c], "<u4").tobytes()
        avg = (sums[zz] / float(c)).astype("<f4") if c else np.zeros(vals.shape[1], "<f4")
        out += avg.tobytes()
    return bytes(out)


def ruleset_full_payload(records, pos) -> bytes:
    """MRSF dump: 24-byte header, the distinct canonical rulesets in first-seen
    order, then per cell (x, y, z: u16, dict_id: u32)."""
    n = len(records)
    raw = np.ascontiguousarray(records).view(np.dtype((np.void, RULESET_FULL_CANONICAL_SIZE)))
    if n:
        _, first, inverse = np.unique(raw, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")      # unique rows by first appearance
        rank = np.empty_like(order)
        rank[order] = np.arange(len(order))
        dict_id = rank[inverse.reshape(-1)]
        dictionary = records[np.sort(first)]
    else:
        dict_id = np.zeros(0, np.int64)
        dictionary = records[:0]
    hdr = RULESET_FULL_MAGIC + np.array([RULESET_FULL_FORMAT_VERSION, 0, len(dictionary), n,
                                         RULESET_FULL_CANONICAL_SIZE], "<u4").tobytes()
    refs = np.zeros(n, np.dtype({"names": ["pos", "id"], "formats": [("<u2", 3), "<u4"], "offsets": [0, 6],
                                 "itemsize": RULESET_FULL_CELL_REF_STRIDE}))
    refs["pos"] = pos
    refs["id"] = dict_id
    return hdr + dictionary.tobytes() + refs.tobytes()


def write_ruleset_layer_dump(rules, pos, grid_z: int, tick: int, out_dir, compression="none", level=3) -> Path:
    Path(out_dir).mkdir(parents=True, exist_ok=True)
    path = binary_file_path(out_dir, tick, "ruleset_layers", compression)
    _write_payload(path, ruleset_layer_payload(canonical_rulesets(rules), np.asarray(pos)[:, 2], grid_z),
                   compression, level)
    return path


def write_ruleset_full_dump(rules, pos, tick: int, out_dir, compression="none", level=3) -> Path:
    Path(out_dir).mkdir(parents=True, exist_ok=True)
    path = binary_file_path(out_dir, tick, "rulesets", compression)
    _write_payload(path, ruleset_full_payload(canonical_rulesets(rules), pos), compression, level)
    return path


# ---------------------------------------------------------------------------
# summary.md (≙ DataLogger::write_summary, crates/marl-output/src/data.rs:467-800)

def summary_text(dims, total_ticks: int, runtime_s: float, summary: dict, z_counts, column: dict,
                 total_div: int, total_death: int, sim) -> str:
    """The Markdown report.  `summary`/`z_counts` as from MarlGpu.cell_summary;
    `column` holds the centre-column reads: oxidant/reductant (all z), carbon,
    waste and light at the surface, middle and deep planes."""
    X, Y, Z = dims
    vox = X * Y * Z
    pop = int(summary["n_cells"])
    pop_pct = pop / vox * 100.0 if vox > 0 else 0.0
    rt = float(np.float32(runtime_s))  # the reference carries the runtime as f32
    tps = total_ticks / rt if rt > 0.0 else 0.0
    L = ["# Run Summary", "", "## Parameters", "",
         f"- Grid: {X}x{Y}x{Z} ({vox} voxels)", f"- Ticks: {total_ticks}",
         f"- Runtime: {rt:.1f}s ({tps:.2f} ticks/sec)", f"- Maintenance rate: {rust_f32(sim.lambda_maintenance)}",
         f"- Boundary sources: oxidant={rust_f32(sim.source_rate_oxidant)}, carbon={rust_f32(sim.source_rate_carbon)}, "
         f"reductant={rust_f32(sim.source_rate_reductant)}", "",
         "## Population", "", f"- Final population: {pop} ({pop_pct:.1f}% capacity)",
         f"- Total divisions: {total_div}", f"- Total deaths: {total_death}"]
    turnover = total_death / total_div if total_div > 0 else 0.0
    L += [f"- Turnover ratio: {turnover:.2f}", "", "## Vertical Zonation", "", "Per-layer cell counts:", "",
          "| z | cells |", "|---|-------|"]
    zc = [int(c) for c in z_counts]
    L += [f"| {z} | {zc[z]} |" for z in range(Z)]
    third, two = Z // 3, 2 * Z // 3
    surf, mid, deep = sum(zc[:third]), sum(zc[third:two]), sum(zc[two:])
    L += ["", f"- Surface zone (z < {third}): {surf} cells", f"- Middle zone ({third} <= z < {two}): {mid} cells",
          f"- Deep zone (z >= {two}): {deep} cells"]
    zmax, zmin = max(surf, mid, deep), min(surf, mid, deep)
    ratio = zmax / zmin if zmin > 0 else (float("inf") if zmax > 0 else 1.0)
    L += [f"- Zonation ratio (max_zone / min_zone): {ratio:.1f}", "", "## Chemical Gradients (center column)", ""]
    ox, red = column["oxidant"], column["reductant"]
    c3, w4, lt = column["carbon"], column["waste"], column["light"]
    zm = Z // 2
    L += [f"- Oxidant: surface={ox[0]:.3f}, mid={ox[zm]:.3f}, deep={ox[Z - 1]:.3f}",
          f"- Reductant: surface={red[0]:.3f}, mid={red[zm]:.3f}, deep={red[Z - 1]:.3f}",
          f"- Carbon: surface={c3[0]:.3f}, mid={c3[1]:.3f}, deep={c3[2]:.3f}",
          f"- Organic waste: surface={w4[0]:.3f}, mid={w4[1]:.3f}, deep={w4[2]:.3f}"]
    f001 = np.float32(0.01)
    ox_pen = next((z for z in range(Z) if np.float32(ox[z]) < f001), Z)
    red_pen = next((z for z in reversed(range(Z)) if np.float32(red[z]) < f001), 0)
    L += [f"- Oxidant penetration depth: z={ox_pen} (first z where oxidant < 0.01)",
          f"- Reductant penetration depth: z={red_pen} (first z from bottom where reductant < 0.01)",
          f"- Light: surface={lt[0]:.3f}, mid={lt[1]:.3f}, deep={lt[2]:.3f}", "", "## Metabolic Diversity", ""]
    if pop > 0:
        L += [f"- Average active reactions: {summary['sum_active_reactions'] / pop:.1f}",
              f"- Average energy: {summary['sum_energy'] / pop:.3f}",
              f"- Average enzyme_A (internal[5]): {summary['sum_enzyme_a'] / pop:.4f}",
              f"- Average enzyme_B (internal[6]): {summary['sum_enzyme_b'] / pop:.4f}"]
    else:
        L += ["- (no living cells — all metrics are zero)"]
    L += ["", "## Assessment", ""]
    if pop_pct > 80.0:
        L.append("- **Grid saturated** — consider larger grid or higher maintenance")
    elif pop_pct < 5.0:
        L.append("- **Population sparse** — dynamics may be resource-limited")
    else:
        L.append(f"- Population at {pop_pct:.1f}% capacity — within normal operating range")
    if ratio > 2.0:
        L.append(f"- **Clear vertical zonation detected** (ratio {ratio:.1f}x)")
    else:
        L.append(f"- Weak or no vertical zonation (ratio {ratio:.1f}x)")
    if ox_pen < Z / 2.0:
        L.append(f"- **Anoxic zone established** — oxidant penetrates to z={ox_pen}")
    else:
        L.append(f"- No clear anoxic zone — oxidant penetrates to z={ox_pen}")
    if 0.0 < turnover < 0.5:
        L.append(f"- Low turnover ({turnover:.2f}) — population is growing")
    elif turnover > 1.5:
        L.append(f"- High turnover ({turnover:.2f}) — population is declining")
    elif total_div > 0:
        L.append(f"- Turnover near unity ({turnover:.2f}) — approximate steady state")
    L.append("")
    return "\n".join(L) + "\n"


def centre_column(gpu) -> dict:
    X, Y, Z = gpu.dims
    cx, cy = X // 2, Y // 2
    q = [(cx, cy, z, 1) for z in range(Z)] + [(cx, cy, z, 2) for z in range(Z)]
    depths = (0, Z // 2, Z - 1)
    q += [(cx, cy, z, 3) for z in depths] + [(cx, cy, z, 4) for z in depths] + [(cx, cy, z, -1) for z in depths]
    v = gpu.sample(q)
    return {"oxidant": v[:Z], "reductant": v[Z:2 * Z], "carbon": v[2 * Z:2 * Z + 3],
            "waste": v[2 * Z + 3:2 * Z + 6], "light": v[2 * Z + 6:2 * Z + 9]}


def write_summary(gpu, out_dir, total_ticks: int, runtime_s: float, total_div: int, total_death: int, sim) -> Path:
    summary, zc = gpu.cell_summary()
    path = Path(out_dir) / "summary.md"
    path.write_text(summary_text(gpu.dims, total_ticks, runtime_s, summary, zc, centre_column(gpu), total_div,
                                 total_death, sim))
    return path
