"""The device-driven tick captured as a hipGraph (marl_tick_graph) against the
host-driven tick (marl_tick) and the oracle: every tick bit-identical —
fields, light and every cell field — including ticks whose cell phase exceeds
the captured schedule budget and is finished by the host-driven path."""
import numpy as np
import pytest

from test_gpu_cells import assert_cells_equal, population

pytestmark = pytest.mark.gpu

F = np.float32


def _compare_runs(pkg, dims, field, cells, rules, sim, n_ticks, limits=None, orc=None, layouts=("auto", "auto"),
                  blocks=(128, 128)):
    seen = {"divisions": 0, "deaths": 0, "host_finished": 0}
    with pkg.MarlGpu(dims) as ge, pkg.MarlGpu(dims) as gg:
        ge.set_cell_layout(layouts[0])
        gg.set_cell_layout(layouts[1])
        ge.kinetics_block, gg.kinetics_block = blocks
        for g in (ge, gg):
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
        if limits:
            gg.set_graph_limits(*limits)
        f, c, r = field, cells, rules
        for t in range(n_ticks):
            se = ge.tick(sim, t, 1)
            sg = gg.tick_graph(sim, t, 1)
            for k in ("divisions", "deaths", "births_added", "n_cells", "levels"):
                assert sg[k] == se[k], (t, k, sg[k], se[k])
            seen["divisions"] += se["divisions"]
            seen["deaths"] += se["deaths"]
            seen["host_finished"] += sg["host_finished"]
            c_e, r_e = ge.download_cells()
            c_g, r_g = gg.download_cells()
            assert_cells_equal(c_g, r_g, c_e, r_e)
            # every byte, padding and reserved words included (both uploads zero them)
            assert c_g.tobytes() == c_e.tobytes() and r_g.tobytes() == r_e.tobytes(), t
            assert np.array_equal(gg.download_field_aos().view(np.uint32), ge.download_field_aos().view(np.uint32)), t
            assert np.array_equal(gg.download_light().view(np.uint32), ge.download_light().view(np.uint32)), t
            if orc is not None:
                f, lt, c, r, _, st = orc.tick(dims, sim, t, f, c, r)
                assert_cells_equal(c_g, r_g, c, r)
                assert np.array_equal(gg.download_field_aos(), f), t
    return seen


@pytest.mark.parametrize("layouts", [("aos", "quads"), ("quads", "aos")])
def test_graph_tick_bit_identical_with_divisions_and_deaths(pkg, orc, layouts):
    """Eager ticks on one field layout of the cell pass, captured ticks on the other."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 4000, 2)
    rng = np.random.default_rng(4)
    field = rng.random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=77, diffusion_substeps=4)
    seen = _compare_runs(pkg, dims, field, cells, rules, sim, 6, orc=orc, layouts=layouts)
    assert seen["divisions"] > 50 and seen["deaths"] > 50, seen
    assert seen["host_finished"] == 0


_LARGE = {}


def _large_population(pkg, dims):
    if dims not in _LARGE:
        _LARGE[dims] = population(pkg._abi, dims, 150_000, 31)
    return _LARGE[dims]


@pytest.mark.parametrize("layouts", [("auto", "aos"), ("aos", "quads")])
def test_graph_tick_overlapped_phases_large_population(pkg, orc, layouts):
    """From 131,072 cells up the captured tick runs the schedule on a second
    stream beside the field phase and writes the AoS view back beside fate."""
    dims = (128, 128, 64)
    cells, rules = _large_population(pkg, dims)
    field = np.random.default_rng(32).random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=81, diffusion_substeps=4)
    seen = _compare_runs(pkg, dims, field, cells, rules, sim, 2, orc=orc, layouts=layouts)
    assert seen["deaths"] > 0 and seen["host_finished"] == 0, seen


@pytest.mark.parametrize("block", [64, 256])
def test_graph_tick_every_kinetics_block(pkg, orc, block):
    """The captured tick's level kernels with the other block sizes."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 4000, 22)
    field = np.random.default_rng(23).random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=79, diffusion_substeps=2)
    _compare_runs(pkg, dims, field, cells, rules, sim, 3, orc=orc, blocks=(128, block))


def test_graph_tick_small_population_one_workgroup(pkg, orc):
    """Populations of at most 256 cells (on the resident quads) run the
    captured cell phase in one workgroup (k_g_cell_phase_small), up to 384
    once there.  A small population that divides and dies for eight ticks
    (240 -> 365 -> 267 cells): every tick equals the eager tick and the oracle."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 240, 41, energy_scale=6.0)
    field = np.random.default_rng(42).random(np.prod(dims) * 12).astype(F) * 4.0
    sim = pkg.SimulationConfig(rng_seed=43, diffusion_substeps=3)
    seen = _compare_runs(pkg, dims, field, cells, rules, sim, 8, orc=orc, layouts=("aos", "auto"))
    print("small population", seen, flush=True)
    assert seen["divisions"] > 100 and seen["deaths"] > 100 and seen["host_finished"] == 0, seen


def test_graph_tick_switches_between_small_and_multi_launch_phases(pkg):
    """The default column with random reactions falls from 4,098 cells to 17
    in its first tick, so the captured tick switches to the one-workgroup
    phase; a population uploaded above 384 cells switches it back (its first
    tick may outgrow the schedule budget fitted to the small one and be
    finished by the host), and a small one again to the one-workgroup phase.
    Each tick equals the eager tick."""
    from marl_ochre_amd import scenarios

    A = pkg._abi
    field, cells, rules, sim = scenarios.winogradsky_column(starter_state=False)
    seen = _compare_runs(pkg, (128, 128, 64), field, cells, rules, sim, 4)
    assert seen["deaths"] > 4000 and seen["host_finished"] == 0, seen
    dims = (64, 32, 24)
    small_c, small_r = population(A, dims, 200, 51)
    big_c, big_r = population(A, dims, 3000, 52)
    fld = np.random.default_rng(53).random(np.prod(dims) * 12).astype(F)
    sim2 = pkg.SimulationConfig(rng_seed=54, diffusion_substeps=2)
    with pkg.MarlGpu(dims) as ge, pkg.MarlGpu(dims) as gg:
        t = 0
        for c, r in ((small_c, small_r), (big_c, big_r), (small_c, small_r)):
            for g in (ge, gg):
                g.upload_field_aos(fld)
                g.upload_cells(c, r)
            for _ in range(2):
                se = ge.tick(sim2, t, 1)
                sg = gg.tick_graph(sim2, t, 1)
                t += 1
                assert sg["host_finished"] == 0 or len(c) > 384, (t, len(c))
                for k in ("divisions", "deaths", "births_added", "n_cells", "levels"):
                    assert sg[k] == se[k], (t, k, sg[k], se[k])
                c_e, r_e = ge.download_cells()
                c_g, r_g = gg.download_cells()
                assert c_g.tobytes() == c_e.tobytes() and r_g.tobytes() == r_e.tobytes(), t
                assert np.array_equal(gg.download_field_aos().view(np.uint32), ge.download_field_aos().view(np.uint32))


def test_graph_tick_host_finishes_over_budget_ticks(pkg):
    """Budgets too small for the schedule force the skip-and-finish path:
    one sweep (relaxation not known to converge) and two levels (too deep)."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 6000, 3)
    rng = np.random.default_rng(5)
    field = rng.random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=9, diffusion_substeps=2)
    for limits in ((1, 64), (64, 2)):
        seen = _compare_runs(pkg, dims, field, cells, rules, sim, 3, limits=limits)
        assert seen["host_finished"] == 3, (limits, seen)


def test_graph_tick_crowded_and_empty(pkg, orc):
    """158 cells in a crowded block and at the walls (birth collisions,
    enclosed cells), and an empty population: small enough that both the
    captured and the host-driven tick run the one-workgroup cell phase, so
    the oracle is the independent check."""
    A = pkg._abi
    from marl_ochre_amd import starters as st

    dims = (32, 16, 12)
    pos = [(x, y, z) for z in range(3, 8) for y in range(4, 9) for x in range(8, 14)]
    pos += [(0, 0, 0), (31, 15, 11), (0, 15, 5), (31, 0, 6), (20, 8, 5), (22, 8, 5), (24, 8, 5), (21, 10, 5)]
    n = len(pos)
    cells = np.zeros(n, A.CELL_DTYPE)
    rules = np.zeros(n, A.RULESET_DTYPE)
    rng = np.random.default_rng(9)
    for i, p in enumerate(pos):
        rules[i] = st.make_ruleset(i % 3)
        cells[i] = st.make_cell(i % 3, p, i)
    cells["internal"] = (rng.random((n, 16)) * 2).astype(F)
    cells["prep_remaining"] = 1
    field = rng.random(np.prod(dims) * 12).astype(F)
    seen = _compare_runs(pkg, dims, field, cells, rules, pkg.SimulationConfig(rng_seed=5), 3, orc=orc)
    assert seen["divisions"] > 10
    empty_c = np.zeros(0, A.CELL_DTYPE)
    empty_r = np.zeros(0, A.RULESET_DTYPE)
    _compare_runs(pkg, dims, field, empty_c, empty_r, pkg.SimulationConfig(), 2)


def test_graph_tick_starter_column(pkg):
    """The reference's default column with its own starter states: 20 ticks."""
    from marl_ochre_amd import scenarios

    field, cells, rules, sim = scenarios.winogradsky_column(starter_state=True)
    _compare_runs(pkg, (128, 128, 64), field, cells, rules, sim, 20)


def test_graph_tick_odd_substeps(pkg):
    """An odd number of stencil launches per tick swaps the field buffers each
    tick: the captured ticks alternate between two graphs, one per parity."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 3000, 8)
    field = np.random.default_rng(6).random(np.prod(dims) * 12).astype(F)
    _compare_runs(pkg, dims, field, cells, rules, pkg.SimulationConfig(rng_seed=3, diffusion_substeps=3), 4)


def test_graph_budget_automatic_after_explicit(pkg):
    """An explicit one-sweep budget host-finishes every tick; (0, 0) switches
    the budget back to automatic: the next tick is still finished by the host
    (the budget it ran with was one sweep), after which the budget fits the
    schedule and the captured ticks finish on the device."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 5000, 11)
    field = np.random.default_rng(12).random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=21, diffusion_substeps=2)
    with pkg.MarlGpu(dims) as ge, pkg.MarlGpu(dims) as gg:
        for g in (ge, gg):
            g.upload_field_aos(field)
            g.upload_cells(cells, rules)
        finished = []
        for t in range(6):
            if t == 0:
                gg.set_graph_limits(1, 64)
            if t == 2:
                gg.set_graph_limits(0, 0)
            se = ge.tick(sim, t, 1)
            sg = gg.tick_graph(sim, t, 1)
            assert se["levels"] > 1
            for k in ("divisions", "deaths", "births_added", "n_cells", "levels"):
                assert sg[k] == se[k], (t, k, sg[k], se[k])
            finished.append(sg["host_finished"])
        c_e, r_e = ge.download_cells()
        c_g, r_g = gg.download_cells()
        assert_cells_equal(c_g, r_g, c_e, r_e)
        assert np.array_equal(gg.download_field_aos().view(np.uint32), ge.download_field_aos().view(np.uint32))
    assert finished == [1, 1, 1, 0, 0, 0], finished
