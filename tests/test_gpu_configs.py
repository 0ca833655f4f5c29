"""The full-size workloads against the oracle, bit for bit.

- stencil + light stress grid, 256x256x128 x 12 species, 5 % blob occupancy:
  a tick's field phase (sources, 8 masked substeps, Beer-Lambert light fused
  into the last substep) against the oracle's boundary sources, substeps and
  light scan (crates/marl-field/src/field.rs:107-250, light.rs:29-59);
- plain kinetics at scale, 1,000,000 cells on 256x256x128: two full ticks
  (crates/marl-sim/src/lib.rs:160-277) against the oracle's tick — field,
  light, every cell and ruleset field, the tick's counters;
- the same workload through the captured graph (phases overlapped on a
  second stream at this population);
- the 32-column uneven ensemble (x = y in {64, 96, 128, 192}, z = x / 2),
  autotuned and advanced in one captured graph per tick: two ticks, each
  column against the oracle's tick with that column's seed.
"""
import os

import numpy as np
import pytest

from test_gpu_cells import assert_cells_equal

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 8)


def _bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("tile", ["autotuned", 11])
def test_stress_grid_field_phase_full_size(pkg, orc, tile):
    from marl_ochre_amd import scenarios

    dims = (256, 256, 128)
    field, occ, sim = scenarios.stencil_stress(dims, seed=0)
    print("stress grid", dims, "tile", tile, flush=True)
    want = orc.diffuse_tick(orc.boundary(field, dims, sim), occ, dims, sim, nthreads=THREADS)
    want_light = orc.light(want, occ, dims, sim)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_occupancy(occ)
        if tile == "autotuned":
            g.autotune(sim)
            g.upload_field_aos(field)
        else:
            g.tile = tile
        z_chunk = g.tile_shape()[1]
        print("tile", g.tile, "z-chunk", z_chunk, "(light fused)" if z_chunk >= dims[2] else "", flush=True)
        assert g.time_field_phase(sim, 1, sources=True) > 0
        got = g.download_field_aos()
        got_light = g.download_light()
    assert _bits_equal(got, want), np.count_nonzero(got != want)
    assert _bits_equal(got_light, want_light), np.count_nonzero(got_light != want_light)
    assert (want_light > 0).any()


def test_kinetics_at_scale_two_ticks(pkg, orc):
    from marl_ochre_amd import scenarios

    dims = (256, 256, 128)
    field, cells, rules, sim = scenarios.kinetics_at_scale(dims, 1_000_000, seed=0)
    assert len(cells) == 1_000_000
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        f, c, r = field, cells, rules
        for t in range(2):
            print("1M-cell tick", t, flush=True)
            f, lt, c, r, ev, st = orc.tick(dims, sim, t, f, c, r, nthreads=THREADS)
            gs = g.tick(sim, t, 1)
            for k in ("divisions", "deaths", "births_added", "n_cells"):
                assert gs[k] == st[k], (t, k, gs[k], st[k])
            assert st["divisions"] + st["deaths"] > 0, st
            cg, rg = g.download_cells()
            assert_cells_equal(cg, rg, c, r)
            assert _bits_equal(g.download_field_aos(), f), t
            assert _bits_equal(g.download_light(), lt), t


def test_kinetics_at_scale_captured_ticks(pkg, orc):
    """The 1,000,000-cell workload through the captured graph, which from
    131,072 cells up runs the schedule on a second stream beside the field
    phase and the transpose back beside fate: two ticks against the oracle."""
    from marl_ochre_amd import scenarios

    dims = (256, 256, 128)
    field, cells, rules, sim = scenarios.kinetics_at_scale(dims, 1_000_000, seed=0)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        f, c, r = field, cells, rules
        for t in range(2):
            print("1M-cell captured tick", t, flush=True)
            f, lt, c, r, ev, st = orc.tick(dims, sim, t, f, c, r, nthreads=THREADS)
            gs = g.tick_graph(sim, t, 1)
            assert gs["host_finished"] == 0
            for k in ("divisions", "deaths", "births_added", "n_cells"):
                assert gs[k] == st[k], (t, k, gs[k], st[k])
            cg, rg = g.download_cells()
            assert_cells_equal(cg, rg, c, r)
            assert _bits_equal(g.download_field_aos(), f), t
            assert _bits_equal(g.download_light(), lt), t


def test_ensemble_32_columns_against_oracle(pkg, orc):
    from marl_ochre_amd import scenarios

    cols = scenarios.ensemble_columns(32, seed=0)
    sizes = {c[0][0] for c in cols}
    assert {128, 192} <= sizes, sizes
    sim = pkg.SimulationConfig()
    seeds = np.array([c[4] for c in cols], np.uint64)
    gs = [pkg.MarlGpu(c[0]) for c in cols]
    try:
        for g, (dims, field, cells, rules, _) in zip(gs, cols):
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
        with pkg.MarlEnsemble(gs) as ens:
            tile = ens.autotune(sim)
            print("ensemble tile", tile, ens.launch_shape(), flush=True)
            state = [(c[1], c[2], c[3]) for c in cols]
            events = 0
            for t in range(2):
                print("ensemble tick", t, flush=True)
                stats, ms, hf = ens.tick(sim, seeds, t, 1)
                assert hf == 0
                for k, (g, (dims, _, _, _, seed)) in enumerate(zip(gs, cols)):
                    f, c, r = state[k]
                    f, lt, c, r, ev, st = orc.tick(dims, sim.replace(rng_seed=int(seed)), t, f, c, r,
                                                   nthreads=THREADS)
                    state[k] = (f, c, r)
                    for key in ("divisions", "deaths", "births_added", "n_cells"):
                        assert stats[k][key] == st[key], (t, k, key, stats[k][key], st[key])
                    events += st["divisions"] + st["deaths"]
                    cg, rg = g.download_cells()
                    assert_cells_equal(cg, rg, c, r)
                    assert _bits_equal(g.download_field_aos(), f), (t, k, dims)
                    assert _bits_equal(g.download_light(), lt), (t, k, dims)
            assert events > 0
    finally:
        for g in gs:
            g.close()
