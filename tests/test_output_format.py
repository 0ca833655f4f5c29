"""Output formats on the host: the known answers of the reference's format
tests (crates/marl-format/src/lib.rs:366-520), the run_meta.json keys of
write_run_meta (crates/marl-output/src/binary_dump.rs:330-429), the CSV and
console layouts of data.rs / stats.rs, and the payload writer's .zst and
temporary-file behaviour. No GPU: a stand-in object supplies the downloads."""
import json

import numpy as np
import pytest


@pytest.fixture(scope="module")
def out(pkg):
    from marl_ochre_amd import output

    return output


def test_format_constants(out):
    # lib.rs:505-520 (test_constants), 497-502 (test_viewer_cell_record_size)
    assert (out.ENDIANNESS, out.FIELD_DTYPE, out.FIELD_LAYOUT, out.CELL_RECORD_STRIDE) == (
        "little", "f32", "z_y_x_species", 25)
    assert out.VIEWER_DTYPE.itemsize == 25
    assert [out.VIEWER_DTYPE.fields[k][1] for k in ("pos", "lineage_id", "starter_type", "energy")] == [0, 12, 20, 21]


def test_field_byte_len_known_answers(out):
    # lib.rs:370-392: basic, zero argument, overflow
    assert out.field_byte_len(128, 128, 64, 12) == 50_331_648
    for args in ((0, 128, 64, 12), (128, 0, 64, 12), (128, 128, 0, 12), (128, 128, 64, 0)):
        assert out.field_byte_len(*args) is None
    m = 2**32 - 1
    assert out.field_byte_len(m, m, m, m) is None


def test_ruleset_layer_stride_known_answers(out):
    # lib.rs:491-495
    assert out.ruleset_layer_value_count(8, 8, 16, 8) == 110
    assert out.ruleset_layer_record_stride(8, 8, 16, 8) == 448


def test_run_meta_keys_and_values(out):
    meta = out.run_meta((128, 128, 64))
    # RunMeta (lib.rs:130-168) flattened ahead of RunMetaFile's own keys (binary_dump.rs:330-360)
    assert list(meta) == [
        "grid_x", "grid_y", "grid_z", "s_ext", "m_int", "field_dtype", "field_layout", "field_byte_len",
        "cell_record_stride", "endianness", "binary_format_version", "binary_compression",
        "binary_compression_level", "write_binary_field", "write_binary_cells", "field_file_pattern",
        "cell_file_pattern", "snapshot_interval", "max_ticks", "field_count", "cell_header_byte_len",
        "cell_count_source", "cell_pos_units", "cell_record_layout", "ruleset_interval", "ruleset_output_mode",
        "ruleset_layer_file_pattern", "ruleset_layer_record_stride", "ruleset_layer_record_layout"]
    assert meta["field_byte_len"] == 50_331_648 and meta["field_count"] == 12_582_912
    assert meta["binary_format_version"] == 2 and meta["binary_compression"] == "none"
    assert meta["binary_compression_level"] == 0
    assert meta["field_file_pattern"] == "tick_<T>.field.bin"
    assert meta["ruleset_layer_record_stride"] == 448
    z = out.run_meta((128, 128, 64), compression="zstd", level=3)
    assert z["cell_file_pattern"] == "tick_<T>.cells.bin.zst" and z["binary_compression_level"] == 3
    assert z["ruleset_layer_file_pattern"] == "tick_<T>.ruleset_layers.bin.zst"
    out.validate_run_meta(meta)
    out.validate_run_meta(z)


@pytest.mark.parametrize("key,value,word", [("endianness", "big", "endianness"),
                                            ("field_layout", "x_y_z_species", "field_layout"),
                                            ("field_byte_len", 1, "field_byte_len"),
                                            ("binary_compression", "brotli", "binary_compression"),
                                            ("binary_format_version", 3, "binary_format_version"),
                                            ("cell_file_pattern", "", "cell_file_pattern")])
def test_run_meta_validate_rejects(out, key, value, word):
    # lib.rs:415-455
    meta = out.run_meta((128, 128, 64))
    meta[key] = value
    with pytest.raises(out.FormatError, match=word):
        out.validate_run_meta(meta)


def test_run_meta_legacy_defaults_validate(out):
    # lib.rs:457-480: a file without the version/compression/pattern keys
    legacy = json.loads("""{"grid_x": 128, "grid_y": 128, "grid_z": 64, "s_ext": 12, "m_int": 8,
        "field_dtype": "f32", "field_layout": "z_y_x_species", "field_byte_len": 50331648,
        "cell_record_stride": 25, "endianness": "little", "write_binary_field": true,
        "write_binary_cells": true}""")
    out.validate_run_meta(legacy)


def test_run_meta_file_is_pretty_json(out, tmp_path):
    p = out.write_run_meta(tmp_path / "run", (64, 32, 16))
    text = p.read_text()
    assert text.startswith('{\n  "grid_x": 64,\n  "grid_y": 32,') and not text.endswith("\n")
    assert json.loads(text) == out.run_meta((64, 32, 16))


def test_tick_row_and_header(out, tmp_path):
    log = out.TickLog(tmp_path, 3)
    log.close()
    assert (tmp_path / "ticks.csv").read_text() == (
        "tick,population,avg_energy,avg_enzyme_a,avg_enzyme_b,avg_active_rxns,divisions_this_tick,"
        "deaths_this_tick,z0_cells,z1_cells,z2_cells\n")
    s = {"n_cells": 3, "sum_energy": 1.0, "sum_enzyme_a": 2.0, "sum_enzyme_b": 0.5, "sum_active_reactions": 7}
    assert out.TickLog.row(12, s, [1, 0, 2], 4, 5) == "12,3,0.333333,0.666667,0.166667,2.3333,4,5,1,0,2\n"
    empty = {"n_cells": 0, "sum_energy": 0.0, "sum_enzyme_a": 0.0, "sum_enzyme_b": 0.0, "sum_active_reactions": 0}
    assert out.TickLog.row(0, empty, [0, 0], 0, 9) == "0,0,0.000000,0.000000,0.000000,0.0000,0,9,0,0\n"


def test_chemistry_csv_layout(out):
    prof = np.zeros((2, 13))
    prof[0, 0] = 1 / 3
    prof[1, 12] = 2.5
    text = out.chemistry_csv(prof)
    lines = text.splitlines()
    assert lines[0] == "z," + ",".join(f"species_{s}_avg" for s in range(12)) + ",light_avg"
    assert lines[1] == "0,0.333333" + ",0.000000" * 12
    assert lines[2] == "1" + ",0.000000" * 12 + ",2.500000"


class _Stub:
    """Downloads of a fixed small state, shaped like MarlGpu's."""

    def __init__(self, out, dims, n_cells):
        rng = np.random.default_rng(3)
        self.dims = dims
        X, Y, Z = dims
        self.field = rng.random((Z, Y, X, 12)).astype(np.float32)
        self.light = rng.random((Z, Y, X)).astype(np.float32)
        rec = np.zeros(n_cells, out.VIEWER_DTYPE)
        rec["pos"] = rng.integers(0, 8, (n_cells, 3))
        rec["lineage_id"] = rng.integers(0, 2**63, n_cells, dtype=np.uint64)
        rec["starter_type"] = rng.integers(0, 3, n_cells)
        rec["energy"] = rng.random(n_cells)
        self.rec = rec
        self.summary = ({"n_cells": n_cells, "sum_energy": 3.0 * n_cells, "sum_enzyme_a": 0.25 * n_cells,
                         "sum_enzyme_b": 0.0, "sum_active_reactions": 3 * n_cells, "z_thirds": [10, 20, 70]},
                        np.arange(Z))

    def download_field_aos(self):
        return self.field.reshape(-1)

    def download_viewer_cells(self):
        return self.rec.tobytes()

    def cell_summary(self):
        return self.summary

    def sample(self, q):
        q = np.asarray(q).reshape(-1, 4)
        return np.array([self.light[z, y, x] if s < 0 else self.field[z, y, x, s] for x, y, z, s in q], np.float32)


@pytest.mark.parametrize("compression", ["none", "zstd"])
def test_snapshot_round_trip(out, tmp_path, compression):
    gpu = _Stub(out, (32, 8, 6), 40)
    run = tmp_path / "run"
    out.write_run_meta(run, gpu.dims, compression=compression)
    fp = out.write_field_dump(gpu, 500, run, compression)
    cp = out.write_cell_dump(gpu, 500, run, compression)
    assert fp.name == "tick_500.field.bin" + (".zst" if compression == "zstd" else "")
    assert cp.name == "tick_500.cells.bin" + (".zst" if compression == "zstd" else "")
    assert sorted(p.name for p in run.iterdir()) == sorted(["run_meta.json", fp.name, cp.name])
    if compression == "zstd":
        assert fp.read_bytes()[:4] == b"\x28\xb5\x2f\xfd"  # zstd frame magic
    meta, field, cells = out.read_snapshot(run, 500)
    assert np.array_equal(field, gpu.field)
    assert cells.tobytes() == gpu.rec.tobytes()
    assert (run / fp.name).stat().st_size > 0


def test_snapshot_reader_checks_sizes(out, tmp_path):
    gpu = _Stub(out, (32, 8, 6), 5)
    out.write_run_meta(tmp_path, gpu.dims)
    out.write_field_dump(gpu, 1, tmp_path)
    (tmp_path / "tick_1.cells.bin").write_bytes(b"\0" * 26)
    with pytest.raises(out.FormatError, match="not divisible by stride 25"):
        out.read_snapshot(tmp_path, 1)
    (tmp_path / "tick_1.field.bin").write_bytes(b"\0" * 8)
    with pytest.raises(out.FormatError, match="field size mismatch"):
        out.read_snapshot(tmp_path, 1)


def test_payload_writer_leaves_no_temp_file(out, tmp_path):
    with pytest.raises(ValueError, match="unknown compression"):
        out._write_payload(tmp_path / "x.bin", b"abc", "brotli", 0)
    assert list(tmp_path.iterdir()) == []


def test_console_lines(out):
    gpu = _Stub(out, (32, 8, 6), 100)
    # print_stats' layout (stats.rs:69-82), values from the stand-in state
    ox_top, ox_mid, red_bot, org_mid = gpu.field[0, 4, 16, 1], gpu.field[3, 4, 16, 1], gpu.field[5, 4, 16, 2], \
        gpu.field[3, 4, 16, 4]
    line = out.stats_line(gpu, 7, 2, 3, 1.0)
    assert line == (f"t=    7 | pop=   100 (top:   10 mid:   20 bot:   70) | E=3.00 enz=0.250 rxn=3.0 | "
                    f"ox={ox_top:.2f}/{ox_mid:.2f} red={red_bot:.2f} org={org_mid:.2f} | 8.0 t/s")
    gpu.summary = ({"n_cells": 0}, np.zeros(6))
    assert out.stats_line(gpu, 42, 1, 9, 1.0) == "t=   42 | EXTINCT | div=1 death=9"
    gpu = _Stub(out, (32, 8, 6), 10)
    text = out.z_profile_text(gpu).splitlines()
    assert text[0] == "  z  cells  light oxidnt reduct carbon organic"
    assert len(text) == 7
    l = gpu.light[2, 4, 16]
    assert text[3].startswith(f"  2      2 {l:>6.3f} ")
