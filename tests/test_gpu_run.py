"""The command-line run (marl-ochre_amd/run.py) end to end on the GPU, against
the same run driven by the oracle: same configuration, priming, seeding and
RNG seed, so every tick's population, field and outputs must agree —
ticks.csv text, the final binary snapshot (field and 25-byte cell records),
the ruleset dumps and the cell CSV (crates/marl-sim/src/lib.rs:34-420 is the
loop both follow)."""
import io

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TICKS = 25


def _config(run, tmp_path, name, extra=()):
    cfg = run.load(["--grid", "64x32x16", "--ticks", str(TICKS), "--stats", "10", "--snapshot", "12", "--seed", "40",
                    "--rng-seed", "5", "--ruleset-interval", "12", "--output", str(tmp_path / name), *extra],
                   cwd=tmp_path)
    cfg.output.write_tick_log = True
    cfg.output.write_csv_snapshots = True
    cfg.output.ruleset_output_mode = "both"
    return cfg


def _primed_field(dims, sim):
    """init_field_boundaries (crates/marl-sim/src/seeding.rs:9-26): the top
    boundary_prime_layers hold oxidant (1) and carbon (3), the bottom ones
    reductant (2)."""
    X, Y, Z = dims
    f = np.zeros((Z, Y, X, 12), np.float32)
    for z in range(sim.boundary_prime_layers):
        f[z, :, :, 1] = sim.boundary_prime_oxidant
        f[z, :, :, 3] = sim.boundary_prime_carbon
    for z in range(Z - sim.boundary_prime_layers, Z):
        f[z, :, :, 2] = sim.boundary_prime_reductant
    return f.reshape(-1)


def _oracle_population(orc, dims, sim, count, seed):
    """run()'s three seed_cells calls (crates/marl-sim/src/lib.rs:59-104) with the
    oracle's seeding: depth bands from the canonical 200-layer depths scaled by
    Z/200 in f32 and cast to u16, widened to lo+1 / lo+3 / lo+3."""
    from marl_ochre_amd import starters as st

    A = orc.A
    f32 = np.float32
    z_scale = f32(dims[2]) / f32(200.0)
    bands = []
    for lo_f, hi_f, widen in ((sim.phototroph_z_lo, sim.phototroph_z_hi, 1.0),
                              (sim.chemolithotroph_z_lo, sim.chemolithotroph_z_hi, 3.0),
                              (sim.anaerobe_z_lo, sim.anaerobe_z_hi, 3.0)):
        lo = int(f32(lo_f) * z_scale)
        bands.append((lo, int(max(f32(hi_f) * z_scale, f32(lo) + f32(widen)))))
    ca = None
    cells, rules = [], []
    for kind, (lo, hi) in enumerate(bands):
        pos, lin, ca = orc.seed_positions(dims, seed, kind, count, sim.seed_margin, lo, hi, ca,
                                          index_base=sum(len(x) for x in cells))
        c = np.zeros(len(pos), A.CELL_DTYPE)
        r = np.zeros(len(pos), A.RULESET_DTYPE)
        for k in range(len(pos)):
            c[k] = st.make_cell(kind, pos[k], int(lin[k]))
            r[k] = st.make_ruleset(kind)
        cells.append(c)
        rules.append(r)
    return np.concatenate(cells), np.concatenate(rules)


def _seq(x):
    x = np.asarray(x, np.float64).reshape(-1)
    return float(np.cumsum(x)[-1]) if x.size else 0.0


def test_run_matches_oracle_run(pkg, orc, tmp_path):
    from marl_ochre_amd import output as out
    from marl_ochre_amd import run

    cfg = _config(run, tmp_path, "gpu")
    printed = io.StringIO()
    assert run.run(cfg, stdout=printed) == 0
    text = printed.getvalue()
    assert "Seeded " in text and "=== FINAL Z-LAYER PROFILE ===" in text
    for t in (0, 10, 20, TICKS - 1):
        assert f"t={t:>5} | " in text, t

    # the same run through the oracle, set up without the product's helpers
    dims = cfg.grid
    sim = cfg.simulation.replace(rng_seed=5)
    f = _primed_field(dims, sim)
    c, r = _oracle_population(orc, dims, sim, 40, 5)
    assert len(c) > 60
    rows, events = [], 0
    for t in range(TICKS):
        f, lt, c, r, _, st = orc.tick(dims, sim, t, f, c, r)
        act = out.active_reactions(r)
        summ = {"n_cells": len(c), "sum_energy": _seq(c["internal"][:, 0]), "sum_enzyme_a": _seq(c["internal"][:, 5]),
                "sum_enzyme_b": _seq(c["internal"][:, 6]), "sum_active_reactions": int(act.sum())}
        zc = np.bincount(c["pos"][:, 2].astype(np.int64), minlength=dims[2])
        rows.append(out.TickLog.row(t, summ, zc, st["divisions"], st["deaths"]).rstrip("\n"))
        events += st["divisions"] + st["deaths"]
    assert events > 0
    gdir = tmp_path / "gpu"
    assert (gdir / "ticks.csv").read_text().splitlines()[1:] == rows

    last = TICKS - 1
    meta, field, recs = out.read_snapshot(gdir, last)   # zstd, the reference's default
    assert meta["binary_compression"] == "zstd" and meta["grid_z"] == dims[2]
    assert np.array_equal(field.reshape(-1).view(np.uint32), f.view(np.uint32))
    want = np.zeros(len(c), out.VIEWER_DTYPE)
    want["pos"] = c["pos"].astype(np.float32)
    want["lineage_id"] = c["lineage_id"]
    want["starter_type"] = c["starter_type"]
    want["energy"] = c["internal"][:, 0]
    assert recs.tobytes() == want.tobytes()
    canon = out.canonical_rulesets(r)
    assert out.read_payload(gdir / f"tick_{last}.rulesets.bin.zst") == out.ruleset_full_payload(canon, c["pos"])
    assert out.read_payload(gdir / f"tick_{last}.ruleset_layers.bin.zst") == \
        out.ruleset_layer_payload(canon, c["pos"][:, 2], dims[2])
    assert (gdir / f"cells_{last}.csv").read_text() == out.cells_csv(c, r)
    for name in ("run_meta.json", "summary.md", "reaction_registry.csv", f"chem_{last}.csv", "tick_0.field.bin.zst",
                 "tick_12.cells.bin.zst", f"reactions_{last}.csv"):
        assert (gdir / name).exists(), name


def test_eager_and_graph_runs_agree(pkg, tmp_path):
    from marl_ochre_amd import run

    a = _config(run, tmp_path, "graph")
    b = _config(run, tmp_path, "eager", ("--eager",))
    assert b.eager and not a.eager
    for cfg in (a, b):
        assert run.run(cfg, stdout=io.StringIO()) == 0
    for name in ("ticks.csv", f"cells_{TICKS - 1}.csv", f"tick_{TICKS - 1}.cells.bin.zst"):
        assert (tmp_path / "graph" / name).read_bytes() == (tmp_path / "eager" / name).read_bytes(), name
