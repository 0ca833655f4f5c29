"""A tick's field phase on the GPU (boundary sources fused into the first
substep; the light scan fused into the last substep when the tile's z-chunk
spans the column, else the separate light kernel) against the oracle's
boundary sources, substeps and light scan (crates/marl-field/src/field.rs:
107-250, crates/marl-field/src/light.rs:29-59): field and light bit-identical
for tiles that fuse and tiles that do not, through marl_time_field_phase, the
host-driven tick and the captured tick."""
import numpy as np
import pytest

from test_gpu_cells import assert_cells_equal, population

pytestmark = pytest.mark.gpu

F = np.float32
DIMS = (64, 48, 24)
# z-chunks 32, 32, 64, 128 span Z = 24 (fused light); 16 and 8 do not
TILES = [0, 5, 8, 11, 13, 15]


def _sim(pkg, substeps):
    # strong absorption so that the intensity crosses the floor inside the column
    return pkg.SimulationConfig(diffusion_substeps=substeps, chemical_absorption=0.6, cell_absorption=0.4,
                                light_floor=2e-3)


def _inputs(seed, dims=DIMS):
    rng = np.random.default_rng(seed)
    nv = dims[0] * dims[1] * dims[2]
    field = rng.random(nv * 12, dtype=F)
    occ = (rng.random(nv) < 0.12).astype(np.uint8)
    return field, occ


@pytest.mark.parametrize("substeps", [1, 4])
@pytest.mark.parametrize("tile", TILES)
def test_field_phase_matches_oracle(pkg, orc, tile, substeps):
    field, occ = _inputs(tile * 10 + substeps)
    sim = _sim(pkg, substeps)
    want = orc.diffuse_tick(orc.boundary(field, DIMS, sim), occ, DIMS, sim)
    want_light = orc.light(want, occ, DIMS, sim)
    assert (want_light == 0).any() and (want_light > 0).any()   # the floor is exercised
    with pkg.MarlGpu(DIMS) as g:
        g.tile = tile
        g.upload_field_aos(field)
        g.upload_occupancy(occ)
        assert g.time_field_phase(sim, 1, sources=True) > 0
        got = g.download_field_aos()
        got_light = g.download_light()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), np.count_nonzero(got != want)
    assert np.array_equal(got_light.view(np.uint32), want_light.view(np.uint32)), \
        np.count_nonzero(got_light != want_light)


@pytest.mark.parametrize("tile", [11, 14])
def test_field_phase_ragged_grid(pkg, orc, tile):
    """X = 96 (a half-filled 64-lane tile), Y = 70 (partial row groups),
    Z = 37: tile 11 (z-chunk 128) fuses the light, tile 14 (z-chunk 16) does
    not."""
    dims = (96, 70, 37)
    field, occ = _inputs(40 + tile, dims)
    sim = _sim(pkg, 3)
    want = orc.diffuse_tick(orc.boundary(field, dims, sim), occ, dims, sim)
    want_light = orc.light(want, occ, dims, sim)
    with pkg.MarlGpu(dims) as g:
        g.tile = tile
        g.upload_field_aos(field)
        g.upload_occupancy(occ)
        g.time_field_phase(sim, 1, sources=True)
        assert np.array_equal(g.download_field_aos().view(np.uint32), want.view(np.uint32))
        assert np.array_equal(g.download_light().view(np.uint32), want_light.view(np.uint32))


def test_field_phase_without_substeps(pkg, orc):
    """diffusion_substeps = 0: boundary sources, then the light scan alone."""
    field, occ = _inputs(3)
    sim = _sim(pkg, 0)
    want = orc.boundary(field, DIMS, sim)
    want_light = orc.light(want, occ, DIMS, sim)
    with pkg.MarlGpu(DIMS) as g:
        g.tile = 11
        g.upload_field_aos(field)
        g.upload_occupancy(occ)
        g.time_field_phase(sim, 1, sources=True)
        assert np.array_equal(g.download_field_aos(), want)
        assert np.array_equal(g.download_light(), want_light)


@pytest.mark.parametrize("tile,substeps", [(11, 4), (13, 4), (11, 3)])
def test_ticks_with_fused_light(pkg, orc, tile, substeps):
    """Whole ticks, host-driven and captured, against the oracle: the fused
    light feeds the cell pass of the same tick (an odd substep count
    alternates the captured tick between two graphs)."""
    A = pkg._abi
    cells, rules = population(A, DIMS, 5000, 21)
    field, _ = _inputs(22)
    sim = _sim(pkg, substeps).replace(rng_seed=31)
    with pkg.MarlGpu(DIMS) as ge, pkg.MarlGpu(DIMS) as gg:
        for g in (ge, gg):
            g.tile = tile
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
        f, c, r = field, cells, rules
        for t in range(3):
            f, lt, c, r, _, st = orc.tick(DIMS, sim, t, f, c, r)
            se = ge.tick(sim, t, 1)
            sg = gg.tick_graph(sim, t, 1)
            for k in ("divisions", "deaths", "births_added", "n_cells"):
                assert se[k] == st[k] and sg[k] == st[k], (t, k, se[k], sg[k], st[k])
            for g in (ge, gg):
                assert np.array_equal(g.download_light().view(np.uint32), lt.view(np.uint32)), t
                assert np.array_equal(g.download_field_aos().view(np.uint32), f.view(np.uint32)), t
                c_g, r_g = g.download_cells()
                assert_cells_equal(c_g, r_g, c, r)
