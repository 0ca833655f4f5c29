"""Ensembles of uneven columns (marl_ensemble_*) against each column ticked
alone with marl_tick: fields, light and every cell field bit-identical after
every tick, for every stencil tile shape; the autotuner leaves fields alone."""
import numpy as np
import pytest

from test_gpu_cells import assert_cells_equal, population

pytestmark = pytest.mark.gpu

F = np.float32


def _columns(pkg):
    """Five columns of different shapes; three with busy random populations
    (divisions and deaths from the first tick), two with starter states."""
    from marl_ochre_amd import scenarios

    A = pkg._abi
    cols = []
    for k, dims in enumerate([(64, 32, 24), (32, 32, 16), (96, 64, 40)]):
        cells, rules = population(A, dims, int(0.08 * np.prod(dims)), 40 + k)
        field = np.random.default_rng(50 + k).random(np.prod(dims) * 12).astype(F)
        cols.append((dims, field, cells, rules, 1000 + k))
    for dims, field, cells, rules, seed in scenarios.ensemble_columns(2, seed=3, sizes=(64, 96),
                                                                       cells_range=(500, 5000)):
        cols.append((dims, field, cells, rules, seed))
    return cols


def _run(pkg, cols, sim, n_ticks, tile=None, per_tile_grid=False, whole=False, chunks=None):
    seeds = np.array([c[4] for c in cols], np.uint64)
    alone = [pkg.MarlGpu(c[0]) for c in cols]
    batched = [pkg.MarlGpu(c[0]) for c in cols]
    try:
        for g, h, (dims, field, cells, rules, _) in zip(alone, batched, cols):
            for x in (g, h):
                x.upload_field_aos(field)
                x.upload_cells(cells, rules)
        events = 0
        with pkg.MarlEnsemble(batched) as ens:
            if tile is not None:
                ens.set_tile(tile)
            ens.set_grid(per_tile_grid)
            ens.set_whole_columns(whole)
            if chunks is not None:
                ens.set_chunks(chunks)
            for t in range(n_ticks):
                stats, ms, hf = ens.tick(sim, seeds, t, 1)
                assert ms > 0 and hf == 0
                for c, (g, h) in enumerate(zip(alone, batched)):
                    se = g.tick(sim.replace(rng_seed=int(seeds[c])), t, 1)
                    for k in ("divisions", "deaths", "births_added", "n_cells"):
                        assert stats[c][k] == se[k], (t, c, k, stats[c][k], se[k])
                    events += se["divisions"] + se["deaths"]
                    ca, ra = g.download_cells()
                    cb, rb = h.download_cells()
                    assert_cells_equal(cb, rb, ca, ra)
                    assert np.array_equal(h.download_field_aos().view(np.uint32),
                                          g.download_field_aos().view(np.uint32)), (t, c)
                    assert np.array_equal(h.download_light().view(np.uint32), g.download_light().view(np.uint32)), (t, c)
        return events
    finally:
        for x in alone + batched:
            x.close()


def test_ensemble_bit_identical_to_single_columns(pkg):
    cols = _columns(pkg)
    events = _run(pkg, cols, pkg.SimulationConfig(diffusion_substeps=4), 4)
    assert events > 50


@pytest.mark.parametrize("tile", [0, 1, 4, 7, 9, 13, 15, 16, 17, 18, 19, 20])
def test_ensemble_every_tile_shape(pkg, tile):
    cols = _columns(pkg)[:3]
    _run(pkg, cols, pkg.SimulationConfig(diffusion_substeps=2), 2, tile=tile)


@pytest.mark.parametrize("tile", [0, 7, 11, 15, 17, 18])
def test_ensemble_one_block_per_tile(pkg, tile):
    """The batched stencil with one block per logical tile instead of the
    persistent grid: same kernel, same results."""
    cols = _columns(pkg)
    _run(pkg, cols, pkg.SimulationConfig(diffusion_substeps=2), 2, tile=tile, per_tile_grid=True)


@pytest.mark.parametrize("tile", [0, 6, 7, 11, 17, 18])
def test_ensemble_balanced_pieces(pkg, tile):
    """Grid mode 2: a persistent grid over balanced whole-depth pieces (the
    last round's tiles cut into z-chunks): same fields, light and cells."""
    cols = _columns(pkg)
    _run(pkg, cols, pkg.SimulationConfig(diffusion_substeps=2), 2, tile=tile, per_tile_grid=2)


@pytest.mark.parametrize("tile,per_tile", [(0, False), (11, True), (15, False), (18, True)])
def test_ensemble_whole_columns_with_fused_light(pkg, tile, per_tile):
    """Every column one z-chunk deep, the light scan carried by the last
    substep of the batched launch: same fields, light and cells."""
    cols = _columns(pkg)
    _run(pkg, cols, pkg.SimulationConfig(diffusion_substeps=2), 2, tile=tile, per_tile_grid=per_tile, whole=True)


@pytest.mark.parametrize("tile,per_tile", [(18, True), (7, False), (0, True)])
def test_ensemble_per_column_chunks(pkg, tile, per_tile):
    """Columns of one batched launch with z-chunks of their own (4, the tile
    shape's, 32, 8, and one longer than the column): same fields, light and
    cells as each column ticked alone."""
    cols = _columns(pkg)
    assert [c[0][2] for c in cols] == [24, 16, 40, 32, 48]
    _run(pkg, cols, pkg.SimulationConfig(diffusion_substeps=4), 2, tile=tile, per_tile_grid=per_tile,
         chunks=[4, 0, 32, 8, 64])


def test_ensemble_chunks_reported_and_checked(pkg):
    cols = _columns(pkg)
    gs = [pkg.MarlGpu(c[0]) for c in cols]
    try:
        with pkg.MarlEnsemble(gs) as ens:
            ens.set_tile(18)  # 16-plane chunks
            ens.set_chunks([4, 0, 32, 8, 64])
            assert ens.chunks() == [4, 16, 32, 8, 48]
            ens.set_whole_columns(True)
            assert ens.chunks() == [24, 16, 40, 32, 48]
            ens.set_whole_columns(False)
            ens.set_chunks(None)
            assert ens.chunks() == [16, 16, 16, 16, 16]
            with pytest.raises(pkg.InvalidInput):
                ens.set_chunks([4, 4])
            with pytest.raises(pkg.InvalidInput):
                ens.set_chunks([4, -1, 4, 4, 4])
    finally:
        for g in gs:
            g.close()


def test_ensemble_rejects_temporal_tiles(pkg):
    """The persistent ensemble kernel runs single-substep tiles only."""
    cols = _columns(pkg)[:2]
    gs = [pkg.MarlGpu(c[0]) for c in cols]
    try:
        with pkg.MarlEnsemble(gs) as ens:
            n = gs[0].lib.marl_stencil_tile_count()
            bad = [t for t in range(n) if gs[0].tile_substeps(t) != 1]
            assert bad
            for t in bad:
                with pytest.raises(pkg.InvalidInput):
                    ens.set_tile(t)
    finally:
        for g in gs:
            g.close()


def test_ensemble_autotune_keeps_fields(pkg):
    cols = _columns(pkg)
    gs = [pkg.MarlGpu(c[0]) for c in cols]
    try:
        for g, (dims, field, cells, rules, _) in zip(gs, cols):
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
        before = [g.download_field_aos() for g in gs]
        with pkg.MarlEnsemble(gs) as ens:
            tile = ens.autotune(pkg.SimulationConfig())
            assert 0 <= tile < pkg.load_library().marl_stencil_tile_count()
            assert ens.launch_shape()["tile"] == tile
            zc = ens.chunks()
            assert all(0 < z <= c[0][2] for z, c in zip(zc, cols)), zc
            assert ens.time_field(pkg.SimulationConfig(), 4) > 0
            # one block per logical tile, or a persistent grid of at most that many blocks
            ens.set_grid(True)
            shape = ens.launch_shape()
            assert shape["grid_blocks"] == shape["stencil_tiles"] > 0
            ens.set_grid(False)
            shape = ens.launch_shape()
            assert 0 < shape["grid_blocks"] <= shape["stencil_tiles"]
            ens.set_grid(2)
            shape = ens.launch_shape()
            assert shape["grid_mode"] == 2 and shape["grid_blocks"] % 8 == 0 and shape["stencil_tiles"] > 0
            with pytest.raises(pkg.InvalidInput):
                ens.set_grid(3)
            with pytest.raises(pkg.InvalidInput):
                ens.tick(pkg.SimulationConfig(diffusion_substeps=3))
        for g, b in zip(gs, before):
            assert np.array_equal(g.download_field_aos(), b)
    finally:
        for g in gs:
            g.close()
