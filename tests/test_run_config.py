"""The command-line run's configuration and writers, on the host.

Config::load precedence and parsing (crates/marl-config/src/lib.rs:255-376,
known answers of its tests at :378-420), the seeding bands of run()
(crates/marl-sim/src/lib.rs:62-105), field priming (seeding.rs:9-26), and the
byte layouts of the cell/reaction CSVs, ruleset dumps (binary_dump.rs:122-328)
and summary.md (data.rs:467-800)."""
import struct

import numpy as np
import pytest


@pytest.fixture(scope="module")
def run(pkg):
    from marl_ochre_amd import run

    return run


@pytest.fixture(scope="module")
def out(pkg):
    from marl_ochre_amd import output

    return output


def test_output_defaults_match_reference(run):
    # lib.rs:259-280 and the test output_config_defaults_match_binary_plan (:382-391)
    o = run.Config().output
    assert (o.max_ticks, o.stats_interval, o.snapshot_interval, o.image_interval, o.seed_count) == (5000, 100, 500,
                                                                                                    500, 30)
    assert o.write_binary_field and o.write_binary_cells
    assert (o.binary_compression, o.binary_compression_level) == ("zstd", 3)
    assert (o.ruleset_interval, o.ruleset_output_mode) == (1000, "off")
    assert not (o.write_tick_log or o.write_csv_snapshots or o.write_ancestry_map or o.write_density_map)
    assert run.load([], cwd="/nonexistent").output.output_dir == "output/run_128x128x64"


def test_help_prints_usage_and_runs_nothing(run, tmp_path, monkeypatch, capsys):
    monkeypatch.chdir(tmp_path)
    assert run.main(["--ticks", "3", "--help"]) == 0
    assert "usage:" in capsys.readouterr().out
    assert list(tmp_path.iterdir()) == []


def test_toml_overrides_then_cli(run, tmp_path):
    (tmp_path / "marl.toml").write_text("""
[simulation]
dt = 2                      # an integer is accepted for an f32 field
diffusion_substeps = 4
d_voxel = [0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.5]
unknown_key = 7             # ignored
[output]
max_ticks = 77
binary_compression = "none"
ruleset_output_mode = "both"
write_tick_log = true
output_dir = "runs/a"
[grid]
x = 64
y = 32
z = 16
[something_else]
k = 1
""")
    cfg = run.load([], cwd=tmp_path)
    assert cfg.simulation.dt == 2.0 and cfg.simulation.diffusion_substeps == 4
    assert cfg.simulation.d_voxel[-1] == 0.5 and cfg.simulation.c_max == 10.0
    assert (cfg.output.max_ticks, cfg.output.binary_compression, cfg.output.ruleset_output_mode) == (77, "none", "both")
    assert cfg.output.write_tick_log and cfg.output.output_dir == "runs/a" and cfg.grid == (64, 32, 16)
    cfg = run.load(["--ticks", "+12", "--stats", "5", "--snapshot", "x", "--output", "runs/b", "--seed", "9",
                    "--ruleset-interval", "3", "--images", "4", "--rng-seed", "11", "--eager", "--gpu-diffusion"],
                   cwd=tmp_path)
    assert (cfg.output.max_ticks, cfg.output.stats_interval, cfg.output.snapshot_interval) == (12, 5, 500)
    assert (cfg.output.output_dir, cfg.output.seed_count, cfg.output.ruleset_interval, cfg.output.image_interval) == (
        "runs/b", 9, 3, 4)
    assert cfg.rng_seed == 11 and cfg.eager and cfg.output.max_ticks == 12
    # a trailing flag without a value is skipped
    assert run.load(["--ticks"], cwd=tmp_path).output.max_ticks == 77


def test_config_flag_and_rejected_files(run, tmp_path):
    (tmp_path / "other.toml").write_text("[output]\nmax_ticks = 5\n")
    assert run.load(["--config", "other.toml"], cwd=tmp_path).output.max_ticks == 5
    bad = {
        "syntax": "[output\nmax_ticks = 5\n",
        "type": "[output]\nmax_ticks = 5\nstats_interval = 1.5\n",
        "negative": "[output]\nmax_ticks = -1\n",
        "enum": "[output]\nbinary_compression = \"brotli\"\n",
        "length": "[simulation]\nd_voxel = [1.0, 2.0]\n",
        "narrow": "[simulation]\ndivision_neighbor_distance = 300\n",
        "bool": "[output]\nwrite_tick_log = 1\n",
    }
    for name, text in bad.items():
        (tmp_path / f"{name}.toml").write_text(text)
        msgs = []
        cfg = run.load(["--config", f"{name}.toml"], cwd=tmp_path, warn=msgs.append)
        assert cfg.output.max_ticks == 5000, name          # the whole file is rejected
        assert msgs and msgs[0].startswith(f"Warning: failed to parse config {name}.toml"), (name, msgs)


def test_seed_bands_and_priming(run, pkg):
    sim = pkg.SimulationConfig()
    # Z = 64: z_scale = 0.32; photo [0, max(0.96, 1)), chemo [25, 41), anaerobe [38, 57)
    assert run.seed_bands((128, 128, 64), sim) == [(0, 1), (25, 41), (38, 57)]
    # Z = 200 reproduces the canonical depths
    assert run.seed_bands((128, 128, 200), sim) == [(0, 3), (80, 130), (120, 180)]
    f = run.primed_field((32, 4, 6), sim).reshape(6, 4, 32, 12)
    assert np.all(f[:2, :, :, 1] == np.float32(0.5)) and np.all(f[:2, :, :, 3] == np.float32(0.3))
    assert np.all(f[4:, :, :, 2] == np.float32(0.5)) and f[2:4].sum() == 0 and f[:2, :, :, 2].sum() == 0


def _population(pkg, n=5):
    from marl_ochre_amd import starters

    A = pkg._abi
    cells = np.zeros(n, A.CELL_DTYPE)
    rules = np.zeros(n, A.RULESET_DTYPE)
    for i in range(n):
        rules[i] = starters.make_ruleset(i % 3)
        cells[i] = starters.make_cell(i % 3, (i, 2 * i, i % 3), 1000 + i)
    return cells, rules


def test_cells_and_reactions_csv(pkg, out):
    cells, rules = _population(pkg)
    cells["internal"][0, 0] = np.float32(1.25)
    cells["age"][0] = 7
    cells["quiescent"][0] = 1
    text = out.cells_csv(cells, rules).splitlines()
    assert text[0] == "x,y,z,energy,enzyme_a,enzyme_b,active_rxns,lineage_id,age,quiescent,starter_type"
    act = out.active_reactions(rules)
    i = cells["internal"][0]
    assert text[1] == f"0,0,0,1.250000,{float(i[5]):.6f},{float(i[6]):.6f},{act[0]},1000,7,1,0"
    reg = out.ReactionRegistry()
    rows = out.reactions_csv(cells, rules, reg).splitlines()
    assert rows[0] == "x,y,z,starter_type,lineage_id,reaction_ids"
    # ids are handed out in first-seen order over cells, then reaction slots
    seen = {}
    for k in range(len(cells)):
        ids = []
        for r in rules["reactions"][k]:
            if abs(r["v_max"]) > np.float32(1e-9):
                ids.append(seen.setdefault((int(r["substrate"]), int(r["product"]), int(r["catalyst"]),
                                            int(r["cofactor"])), len(seen)))
        assert rows[k + 1].split(",")[-1] == ";".join(map(str, ids))
    assert reg.count() == len(seen)


def test_canonical_ruleset_layout(pkg, out):
    cells, rules = _population(pkg, 2)
    rec = out.canonical_rulesets(rules)
    b = rec[1].tobytes()
    assert len(b) == 536
    r = rules[1]
    # receptor 0 at 0, transport 0 at 96 (f32, f32, u8 ext, u8 int), reaction 0 at 176, effector 0 at 432,
    # fate at 512, hgt/mutation at 528 (binary_dump.rs:188-240)
    assert struct.unpack_from("<3f", b, 0) == tuple(np.float32([r["receptors"][0]["k_half"], r["receptors"][0]["n_hill"],
                                                                 r["receptors"][0]["gain"]]).tolist())
    t = r["transport"][0]
    assert struct.unpack_from("<2f2B", b, 96) == (float(np.float32(t["uptake_rate"])),
                                                  float(np.float32(t["secrete_rate"])), int(t["ext_species"]),
                                                  int(t["int_species"]))
    x = r["reactions"][3]
    assert struct.unpack_from("<4B3f", b, 176 + 3 * 16) == (int(x["substrate"]), int(x["product"]), int(x["catalyst"]),
                                                           int(x["cofactor"]), float(x["k_m"]), float(x["v_max"]),
                                                           float(x["k_cat"]))
    e = r["effectors"][7]
    assert struct.unpack_from("<2f2B", b, 432 + 70) == (float(e["threshold"]), float(e["rate"]), int(e["int_species"]),
                                                        int(e["ext_species"]))
    assert struct.unpack_from("<4f", b, 512) == (float(r["division_energy"]), float(r["death_energy"]),
                                                 float(r["quiescence_energy"]), float(r["division_prep_ticks"]))
    assert struct.unpack_from("<2f", b, 528) == (float(r["hgt_propensity"]), float(r["mutation_rate"]))


def test_ruleset_dumps(pkg, out, tmp_path):
    cells, rules = _population(pkg, 6)
    rules[3] = rules[0]                 # a repeat of an earlier ruleset
    rules["mutation_rate"][5] = 0.25    # a ruleset that differs only in one value
    pos = cells["pos"].copy()
    pos[:, 2] = [0, 0, 2, 2, 2, 1]
    lay = out.ruleset_layer_payload(out.canonical_rulesets(rules), pos[:, 2], 4)
    stride = out.ruleset_layer_record_stride()
    assert stride == 448 and len(lay) == 4 * stride
    z, res, cnt = struct.unpack_from("<HHI", lay, 2 * stride)
    assert (z, res, cnt) == (2, 0, 3)
    vals = np.frombuffer(lay, "<f4", count=110, offset=2 * stride + 8)
    want_mut = np.float32((np.float64(rules["mutation_rate"][2]) + np.float64(rules["mutation_rate"][3])
                           + np.float64(rules["mutation_rate"][4])) / 3.0)
    assert vals[-1] == want_mut
    assert struct.unpack_from("<HHI", lay, 3 * stride) == (3, 0, 0)
    assert not np.frombuffer(lay, "<f4", count=110, offset=3 * stride + 8).any()
    full = out.ruleset_full_payload(out.canonical_rulesets(rules), pos)
    magic, ver, flags, ndict, ncell, size = struct.unpack_from("<4s5I", full, 0)
    assert (magic, ver, flags, ncell, size) == (b"MRSF", 1, 0, 6, 536)
    assert ndict == len({r.tobytes() for r in out.canonical_rulesets(rules)})
    assert len(full) == 24 + ndict * 536 + 6 * 10       # the check of scripts/check_binary_dump.py
    ids = [struct.unpack_from("<3HI", full, 24 + ndict * 536 + 10 * k) for k in range(6)]
    # first-seen dictionary order: rulesets 3 and 4 repeat 0 and 1 (same starters), 5 is new
    assert [i[3] for i in ids] == [0, 1, 2, 0, 1, 3]
    assert [i[:3] for i in ids] == [tuple(int(v) for v in p) for p in pos]
    canon = out.canonical_rulesets(rules)
    assert full[24:24 + 536] == canon[0].tobytes() and full[24 + 3 * 536:24 + 4 * 536] == canon[5].tobytes()
    p1 = out.write_ruleset_full_dump(rules, pos, 9, tmp_path, "zstd")
    assert p1.name == "tick_9.rulesets.bin.zst" and out.read_payload(p1) == full
    assert out.write_ruleset_layer_dump(rules, pos, 4, 9, tmp_path).read_bytes() == lay


def test_summary_text(pkg, out):
    sim = pkg.SimulationConfig()
    s = {"n_cells": 10, "sum_energy": 5.0, "sum_enzyme_a": 1.0, "sum_enzyme_b": 2.0, "sum_active_reactions": 33}
    Z = 6
    col = {"oxidant": np.float32([0.5, 0.2, 0.009, 0.0, 0.0, 0.0]), "reductant": np.float32([0.0, 0.0, 0.0, 0.3, 0.5, 0.6]),
           "carbon": np.float32([0.1, 0.2, 0.3]), "waste": np.float32([0, 0, 0]), "light": np.float32([1, 0.5, 0.25])}
    t = out.summary_text((32, 4, Z), 100, 2.0, s, [4, 0, 0, 0, 3, 3], col, 8, 2, sim)
    assert t.startswith("# Run Summary\n\n## Parameters\n\n- Grid: 32x4x6 (768 voxels)\n- Ticks: 100\n"
                        "- Runtime: 2.0s (50.00 ticks/sec)\n- Maintenance rate: 0.12\n"
                        "- Boundary sources: oxidant=0.4, carbon=0.15, reductant=0.5\n")
    assert "- Final population: 10 (1.3% capacity)\n" in t and "- Turnover ratio: 0.25\n" in t
    assert "- Surface zone (z < 2): 4 cells\n- Middle zone (2 <= z < 4): 0 cells\n- Deep zone (z >= 4): 6 cells\n" in t
    assert "- Zonation ratio (max_zone / min_zone): inf\n" in t
    assert "- Oxidant penetration depth: z=2 " in t and "- Reductant penetration depth: z=2 " in t
    assert "- Average active reactions: 3.3\n- Average energy: 0.500\n" in t
    assert "- **Population sparse**" in t and "- **Clear vertical zonation detected** (ratio infx)" in t
    assert "- **Anoxic zone established** — oxidant penetrates to z=2" in t and "- Low turnover (0.25)" in t
