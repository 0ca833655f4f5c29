"""The output reductions and packing on the GPU (marl_cell_summary,
marl_chemistry_profile, marl_download_viewer_cells, marl_sample) against the
same quantities computed in numpy the reference's way from the downloaded
state: FP64 sums in cell / voxel order as log_tick and snapshot_chemistry take
them (crates/marl-output/src/data.rs:195-335), integer counts exactly, viewer
records byte for byte (binary_dump.rs:34-41), and the written snapshot read back."""
import numpy as np
import pytest

from test_gpu_cells import population

pytestmark = pytest.mark.gpu

F = np.float32


def _seq_sum(x):
    """FP64 sum taken one value after another (cumsum is sequential)."""
    x = np.asarray(x, np.float64).reshape(-1)
    return float(np.cumsum(x)[-1]) if x.size else 0.0


def _expected_summary(cells, rules, Z):
    active = (np.abs(rules["reactions"]["v_max"]) > F(1e-9)).sum(axis=1)
    z = cells["pos"][:, 2].astype(np.int64)
    zt = Z // 3
    return {
        "n_cells": len(cells),
        "sum_energy": _seq_sum(cells["internal"][:, 0]),
        "sum_enzyme_a": _seq_sum(cells["internal"][:, 5]),
        "sum_enzyme_b": _seq_sum(cells["internal"][:, 6]),
        "sum_active_reactions": int(active.sum()),
        "z_thirds": [int((z < zt).sum()), int(((z >= zt) & (z < 2 * zt)).sum()), int((z >= 2 * zt).sum())],
    }, np.bincount(z, minlength=Z)


def _check_summary(g, dims):
    cells, rules = g.download_cells()
    want, wz = _expected_summary(cells, rules, dims[2])
    got, gz = g.cell_summary()
    for k in ("n_cells", "sum_active_reactions", "z_thirds"):
        assert got[k] == want[k], (k, got[k], want[k])
    for k in ("sum_energy", "sum_enzyme_a", "sum_enzyme_b"):
        assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-9), k
    assert np.array_equal(gz, wz)
    return cells, rules, got, gz


def _viewer_bytes(pkg, cells):
    rec = np.zeros(len(cells), pkg.output.VIEWER_DTYPE)
    rec["pos"] = cells["pos"].astype(F)
    rec["lineage_id"] = cells["lineage_id"]
    rec["starter_type"] = cells["starter_type"]
    rec["energy"] = cells["internal"][:, 0]
    return rec.tobytes()


def test_summary_profile_records_after_ticks(pkg, tmp_path):
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 5000, 21)
    field = np.random.default_rng(2).random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=11, diffusion_substeps=4)
    out = pkg.output
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        log = out.TickLog(tmp_path, dims[2])
        for t in range(3):
            st = g.tick(sim, t, 1)
            c, r, summ, zc = _check_summary(g, dims)
            want, wz = _expected_summary(c, r, dims[2])
            assert out.TickLog.row(t, summ, zc, st["divisions"], st["deaths"]) == \
                out.TickLog.row(t, want, wz, st["divisions"], st["deaths"])
            log.log_tick(g, t, st["divisions"], st["deaths"])
        log.close()
        assert len((tmp_path / "ticks.csv").read_text().splitlines()) == 4

        # chemistry profile: XY-plane means in FP64, voxel order
        X, Y, Z = dims
        f = g.download_field_aos().reshape(Z, Y, X, 12)
        light = g.download_light().reshape(Z, Y, X)
        prof = g.chemistry_profile()
        want = np.zeros((Z, 13))
        for z in range(Z):
            for s in range(12):
                want[z, s] = _seq_sum(f[z, :, :, s]) / (X * Y)
            want[z, 12] = _seq_sum(light[z]) / (X * Y)
        np.testing.assert_allclose(prof, want, rtol=1e-12, atol=1e-15)
        assert out.chemistry_csv(prof) == out.chemistry_csv(want)

        # viewer records, byte for byte
        c, _ = g.download_cells()
        assert g.download_viewer_cells() == _viewer_bytes(pkg, c)

        # point samples, including light and the grid corners
        rng = np.random.default_rng(8)
        q = np.stack([rng.integers(0, X, 200), rng.integers(0, Y, 200), rng.integers(0, Z, 200),
                      rng.integers(-1, 12, 200)], 1)
        q[:4] = [[0, 0, 0, -1], [X - 1, Y - 1, Z - 1, 11], [X - 1, 0, Z - 1, -1], [0, Y - 1, 0, 0]]
        s = g.sample(q)
        ref = np.array([light[z, y, x] if sp < 0 else f[z, y, x, sp] for x, y, z, sp in q], F)
        assert np.array_equal(s.view(np.uint32), ref.view(np.uint32))
        with pytest.raises(pkg.InvalidInput):
            g.sample([[X, 0, 0, 0]])
        with pytest.raises(pkg.InvalidInput):
            g.sample([[0, 0, 0, 12]])

        # console lines from the GPU state
        line = out.stats_line(g, 2, 1, 1, 1.0)
        assert line.startswith("t=    2 | pop=")
        assert len(out.z_profile_text(g).splitlines()) == Z + 1

        # snapshot files, both compressions, read back
        for comp in ("none", "zstd"):
            run = tmp_path / comp
            out.write_run_meta(run, dims, compression=comp)
            out.write_field_dump(g, 3, run, comp)
            out.write_cell_dump(g, 3, run, comp)
            meta, rf, rc = out.read_snapshot(run, 3)
            assert np.array_equal(rf, f) and rc.tobytes() == _viewer_bytes(pkg, c)
            assert len(rc) == len(c)


def test_empty_population(pkg):
    A = pkg._abi
    dims = (32, 8, 6)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(np.ones(np.prod(dims) * 12, F))
        g.upload_cells(np.zeros(0, A.CELL_DTYPE), np.zeros(0, A.RULESET_DTYPE))
        s, zc = g.cell_summary()
        assert s["n_cells"] == 0 and s["sum_energy"] == 0.0 and s["z_thirds"] == [0, 0, 0]
        assert not zc.any()
        assert g.download_viewer_cells() == b""
        assert pkg.output.stats_line(g, 0, 0, 4, 1.0) == "t=    0 | EXTINCT | div=0 death=4"
        np.testing.assert_array_equal(g.chemistry_profile()[:, :12], 1.0)


def test_deep_column_global_histogram(pkg):
    """Z above the 2048 layers the summary kernel counts in LDS: the per-layer
    counts go straight to global memory."""
    A = pkg._abi
    dims = (32, 4, 2304)
    cells, rules = population(A, dims, 3000, 5)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(np.zeros(np.prod(dims) * 12, F))
        g.upload_cells(cells, rules)
        _check_summary(g, dims)
        assert g.download_viewer_cells() == _viewer_bytes(pkg, g.download_cells()[0])
