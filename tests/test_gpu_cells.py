"""GPU parity of the resident cell path through the C ABI.

The level-scheduled cell pass, fate, division with mutation and compaction are
compared with the oracle's sequential restatement of crates/marl-sim/src/lib.rs
:160-277 on the same seed: integer state (positions, order, events, lineages,
ages, prep counters, quiescence) and every f32 value bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = np.float32


def population(A, dims, n, seed, energy_scale=2.0):
    from marl_ochre_amd import scenarios
    from marl_ochre_amd import starters as st

    rng = np.random.default_rng(seed)
    nv = dims[0] * dims[1] * dims[2]
    vox = rng.choice(nv, n, replace=False)
    cells = np.zeros(n, A.CELL_DTYPE)
    rules = np.zeros(n, A.RULESET_DTYPE)
    for i, v in enumerate(vox):
        kind = int(rng.integers(0, 3))
        rules[i] = st.make_ruleset(kind)
        cells[i] = st.make_cell(kind, (v % dims[0], (v // dims[0]) % dims[1], v // (dims[0] * dims[1])),
                                int(rng.integers(0, 2**62)))
    cells["internal"] = (rng.random((n, 16)) * energy_scale).astype(F)
    cells["prep_remaining"] = rng.integers(0, 3, n)
    cells["quiescent"] = rng.integers(0, 2, n)
    half = n // 2
    scenarios.random_reactions(rules[:half], rng, (1e-3, 1.0), (1e-3, 1.0), light_p=0.2)
    rules["mutation_rate"] = 0.3
    return cells, rules


def _leaves(a, prefix=""):
    """(name, array) for every scalar field of a structured array, recursively."""
    if a.dtype.names is None:
        yield prefix, a
        return
    for n in a.dtype.names:
        yield from _leaves(a[n], f"{prefix}.{n}" if prefix else n)


def assert_cells_equal(got_c, got_r, want_c, want_r):
    """Field by field (explicit padding bytes are not state)."""
    assert len(got_c) == len(want_c), (len(got_c), len(want_c))
    if len(got_c) == 0:
        return
    bad_fields = []
    for (name, a), (_, b) in zip(_leaves(got_c), _leaves(want_c)):
        bad = np.flatnonzero(np.any((a != b).reshape(len(a), -1), axis=1))
        if bad.size:
            bad_fields.append(f"cell.{name} at {bad[:8]} got {a[bad[:3]]} want {b[bad[:3]]}")
    for (name, a), (_, b) in zip(_leaves(got_r), _leaves(want_r)):
        bad = np.flatnonzero(np.any((a != b).reshape(len(a), -1), axis=1))
        if bad.size:
            bad_fields.append(f"rules.{name} at {bad[:8]} ({bad.size} cells) got {a[bad[:2]]} want {b[bad[:2]]}")
    assert not bad_fields, "\n".join(bad_fields)


LAYOUTS = pytest.mark.parametrize("layout", ["aos", "quads"])


@LAYOUTS
def test_cell_pass_bit_identical(pkg, orc, layout):
    """The pass on the transposed AoS view and on the resident quads in place."""
    A = pkg._abi
    dims = (64, 48, 32)
    cells, rules = population(A, dims, 12000, 1)
    rng = np.random.default_rng(3)
    nv = np.prod(dims)
    field = rng.random(nv * 12).astype(F)
    light = rng.random(nv).astype(F)
    sim = pkg.SimulationConfig()
    f_w, c_w, ev_w = orc.cell_pass(dims, sim, field, light, cells, rules)
    with pkg.MarlGpu(dims) as g:
        g.set_cell_layout(layout)
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        g.upload_light(light)
        ev, levels = g.cell_pass(sim)
        c_g, r_g = g.download_cells()
        f_g = g.download_field_aos()
    print("levels", levels)
    assert levels >= 3
    assert np.array_equal(ev, ev_w)
    assert_cells_equal(c_g, r_g, c_w, rules)
    assert np.array_equal(f_g, f_w), np.count_nonzero(f_g != f_w)


def _run_ticks(pkg, orc, dims, field, cells, rules, sim, n_ticks, check_every=1, layout="auto", block=None):
    stats_seen = {"divisions": 0, "deaths": 0}
    with pkg.MarlGpu(dims) as g:
        g.set_cell_layout(layout)
        if block is not None:
            g.kinetics_block = block
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        f, c, r = field, cells, rules
        for t in range(n_ticks):
            f, lt, c, r, ev, st = orc.tick(dims, sim, t, f, c, r)
            gs = g.tick(sim, t, 1)
            for k in ("divisions", "deaths", "births_added", "n_cells"):
                assert gs[k] == st[k], (t, k, gs[k], st[k])
            stats_seen["divisions"] += st["divisions"]
            stats_seen["deaths"] += st["deaths"]
            if (t + 1) % check_every == 0 or t == n_ticks - 1:
                c_g, r_g = g.download_cells()
                assert_cells_equal(c_g, r_g, c, r)
                f_g = g.download_field_aos()
                assert np.array_equal(f_g, f), (t, np.count_nonzero(f_g != f))
                assert np.array_equal(g.download_light(), lt), t
    return stats_seen


@LAYOUTS
def test_ticks_with_divisions_deaths_and_mutation_bit_identical(pkg, orc, layout):
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 4000, 2)
    rng = np.random.default_rng(4)
    field = rng.random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=77, diffusion_substeps=4)
    seen = _run_ticks(pkg, orc, dims, field, cells, rules, sim, 6, layout=layout)
    assert seen["divisions"] > 50 and seen["deaths"] > 50, seen


@pytest.mark.parametrize("case", ["temporal_even", "temporal_odd", "split_species", "phase_budget"])
def test_ticks_on_the_view_for_every_kind_of_last_substep(pkg, orc, case):
    """The pass on the AoS view, which the tick's last substep writes when it is a
    single-substep launch (temporal tile with an odd count, a tile whose blocks hold
    some of the species) and which is transposed from the quads otherwise (a
    temporal tile with an even count, the per-phase mass budget): each against the
    oracle over ticks with divisions and deaths."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 4000, 21)
    field = np.random.default_rng(22).random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=79, diffusion_substeps=5 if case == "temporal_odd" else 4)
    with pkg.MarlGpu(dims) as g:
        n = g.lib.marl_stencil_tile_count()
        if case.startswith("temporal"):
            g.tile = next(t for t in range(n) if g.tile_substeps(t) == 2)
        elif case == "split_species":
            g.tile = next(t for t in range(n)
                          if g.tile_substeps(t) == 1 and g.tile_split(t)[0] * (g.tile_split(t)[1] // 4) < 12)
        else:
            g.set_phase_budget(True)
        g.set_cell_layout("aos")
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        f, c, r = field, cells, rules
        seen = 0
        for t in range(3):
            f, lt, c, r, ev, st = orc.tick(dims, sim, t, f, c, r)
            gs = g.tick(sim, t, 1)
            assert gs["n_cells"] == st["n_cells"], (t, gs["n_cells"], st["n_cells"])
            seen += st["divisions"] + st["deaths"]
            c_g, r_g = g.download_cells()
            assert_cells_equal(c_g, r_g, c, r)
            f_g = g.download_field_aos()
            assert np.array_equal(f_g, f), (t, np.count_nonzero(f_g != f))
            assert np.array_equal(g.download_light(), lt), t
    assert seen > 50


@pytest.mark.parametrize("block", [64, 256])
@LAYOUTS
def test_ticks_every_kinetics_block(pkg, orc, layout, block):
    """The plain pass with each block size the kinetics autotuner chooses from."""
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 4000, 12)
    field = np.random.default_rng(13).random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=78, diffusion_substeps=2)
    _run_ticks(pkg, orc, dims, field, cells, rules, sim, 3, layout=layout, block=block)


def test_kinetics_autotune_keeps_state(pkg):
    A = pkg._abi
    dims = (64, 48, 32)
    cells, rules = population(A, dims, 12000, 14)
    field = np.random.default_rng(15).random(np.prod(dims) * 12).astype(F)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        b = g.autotune_kinetics(pkg.SimulationConfig())
        assert b in (64, 128, 256) and g.kinetics_block == b
        assert np.array_equal(g.download_field_aos(), field)
        c2, r2 = g.download_cells()
        assert_cells_equal(c2, r2, cells, rules)
        with pytest.raises(pkg.InvalidInput):
            g.kinetics_block = 96


@LAYOUTS
def test_crowded_population_and_walls(pkg, orc, layout):
    """Dense blocks (enclosed cells lose their deltas), cells on every wall,
    birth collisions (two parents picking one voxel)."""
    A = pkg._abi
    dims = (32, 16, 12)
    from marl_ochre_amd import starters as st

    pos = []
    for z in range(3, 8):
        for y in range(4, 9):
            for x in range(8, 14):
                pos.append((x, y, z))
    pos += [(0, 0, 0), (31, 15, 11), (0, 15, 5), (31, 0, 6), (16, 0, 0), (16, 15, 11)]
    pos += [(20, 8, 5), (22, 8, 5), (24, 8, 5), (21, 10, 5)]   # shared distance-2 targets
    n = len(pos)
    cells = np.zeros(n, A.CELL_DTYPE)
    rules = np.zeros(n, A.RULESET_DTYPE)
    rng = np.random.default_rng(9)
    for i, p in enumerate(pos):
        rules[i] = st.make_ruleset(i % 3)
        cells[i] = st.make_cell(i % 3, p, i)
    cells["internal"] = (rng.random((n, 16)) * 2).astype(F)
    cells["prep_remaining"] = 1   # everyone divides on tick 0 unless it dies
    field = rng.random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=5)
    seen = _run_ticks(pkg, orc, dims, field, cells, rules, sim, 3, layout=layout)
    assert seen["divisions"] > 10


def test_empty_population_tick(pkg, orc):
    A = pkg._abi
    dims = (32, 8, 8)
    rng = np.random.default_rng(1)
    field = rng.random(np.prod(dims) * 12).astype(F)
    cells = np.zeros(0, A.CELL_DTYPE)
    rules = np.zeros(0, A.RULESET_DTYPE)
    _run_ticks(pkg, orc, dims, field, cells, rules, pkg.SimulationConfig(), 2)


def test_winogradsky_column_30_ticks_bit_identical(pkg, orc):
    """The default 128x128x64 column with ~4,096 seeded cells."""
    from marl_ochre_amd import scenarios

    field, cells, rules, sim = scenarios.winogradsky_column()
    _run_ticks(pkg, orc, (128, 128, 64), field, cells, rules, sim, 30, check_every=10)


def test_mass_budget_matches_numpy_f64(pkg):
    A = pkg._abi
    dims = (64, 32, 16)
    cells, rules = population(A, dims, 3000, 5)
    rng = np.random.default_rng(6)
    field = rng.random(np.prod(dims) * 12).astype(F)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        mb = g.mass_budget()
    occ = np.zeros(np.prod(dims), bool)
    for c in cells:
        x, y, z = (int(v) for v in c["pos"])
        occ[(z * dims[1] + y) * dims[0] + x] = True
    f = field.reshape(-1, 12).astype(np.float64)
    assert np.allclose(mb[:12], f[~occ].sum(0), rtol=1e-12)
    assert np.allclose(mb[12:24], f[occ].sum(0), rtol=1e-12)
    assert np.allclose(mb[24:40], cells["internal"].astype(np.float64).sum(0), rtol=1e-12)


def test_phase_budget_matches_numpy_f64_at_each_phase(pkg, orc):
    """marl_tick's per-phase FP64 budgets against numpy FP64 sums of the
    oracle's state at the same points of the tick: after the occupancy
    rebuild, after the field phase, after the cell pass, after fate."""
    A = pkg._abi
    dims = (64, 32, 16)
    cells, rules = population(A, dims, 3000, 31)
    rng = np.random.default_rng(32)
    field = rng.random(np.prod(dims) * 12).astype(F)
    sim = pkg.SimulationConfig(rng_seed=33, diffusion_substeps=2)

    def budget(f, c):
        occ = np.zeros(np.prod(dims), bool)
        for x, y, z in c["pos"].astype(np.int64):
            occ[(z * dims[1] + y) * dims[0] + x] = True
        f = f.reshape(-1, 12).astype(np.float64)
        return np.concatenate([f[~occ].sum(0), f[occ].sum(0), c["internal"].astype(np.float64).sum(0)]), occ

    b0, occ = budget(field, cells)
    f1 = orc.diffuse_tick(orc.boundary(field, dims, sim), occ.astype(np.uint8), dims, sim)
    lt = orc.light(f1, occ.astype(np.uint8), dims, sim)
    b1, _ = budget(f1, cells)
    f2, c2, _ = orc.cell_pass(dims, sim, f1, lt, cells, rules)
    b2, _ = budget(f2, c2)
    f3, _, c3, _, _, st = orc.tick(dims, sim, 0, field, cells, rules)
    b3, _ = budget(f3, c3)
    assert st["divisions"] + st["deaths"] > 0
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        g.set_phase_budget(True)
        g.tick(sim, 0, 1)
        got = g.tick_phase_budget()
    for k, want in enumerate((b0, b1, b2, b3)):
        assert np.allclose(got[k], want, rtol=1e-12, atol=1e-9), (k, np.abs(got[k] - want).max())
    # the field phase adds the boundary sources: species 1, 2 and 3 gain, the
    # cells' internals are untouched; the cell pass moves matter, fate removes it
    assert np.array_equal(got[1, 24:], got[0, 24:])
