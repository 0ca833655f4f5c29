"""Synthetic workloads (setup only, not timed).

Each generator is seeded and documents how the workload description maps onto
the reference model (crates/marl-config/src/lib.rs, crates/marl-field/src/field.rs).
"""
from __future__ import annotations

import numpy as np

from .config import SimulationConfig

F32 = np.float32


def deposit_species7(factor: np.ndarray, sim: SimulationConfig) -> np.ndarray:
    """Species-7 value whose niche factor 1 - a*s7/(k+s7) (field.rs:141-143) equals `factor`.

    Solving f = 1 - a*s/(k+s) gives s = k*(1-f)/(f-(1-a)); factor must exceed 1-a.
    Keeps the reference's semantics (the factor stays a function of species 7)
    at zero extra bytes per voxel.
    """
    a, k = np.float64(sim.alpha_eps), np.float64(sim.k_eps)
    f = factor.astype(np.float64)
    return (k * (1.0 - f) / (f - (1.0 - a))).astype(F32)


def stencil_stress(dims=(256, 256, 128), seed=0, occ_frac=0.05, sigma=1.5):
    """Stencil + light stress grid.

    - 12 species uniform in [0,1] (species 7 set from the deposit factor, below);
    - occupancy: Gaussian-smoothed white noise (sigma voxels, periodic) thresholded
      at its (1 - occ_frac) quantile -> occ_frac of voxels in blobs;
    - deposit diffusivity factor uniform in [0.201, 1] per voxel, realised through
      species 7 (deposit_species7);
    - D_s uniform in [0.05, 0.16] per species, 8 substeps per tick, dt = 1;
    - attenuation: exp(-0.5) per occupied voxel (cell_absorption = 0.5) and
      exp(-0.02 * s4) per voxel (chemical_absorption = 0.02; mean s4 = 0.5 gives
      0.01 per empty voxel).
    Returns (field_aos f32, occupancy u8, SimulationConfig).
    """
    from scipy.ndimage import gaussian_filter

    X, Y, Z = dims
    nv = X * Y * Z
    rng = np.random.Generator(np.random.Philox(seed))
    sim = SimulationConfig(diffusion_substeps=8, cell_absorption=0.5, chemical_absorption=0.02)
    sim.d_voxel = [float(v) for v in rng.uniform(0.05, 0.16, 12).astype(F32)]
    field = rng.random((nv, 12), dtype=F32)
    noise = rng.standard_normal((Z, Y, X), dtype=F32)
    sm = gaussian_filter(noise, sigma=sigma, mode="wrap")
    thr = np.quantile(sm, 1.0 - occ_frac)
    occ = (sm > thr).astype(np.uint8).reshape(-1)
    factor = rng.uniform(0.201, 1.0, nv)
    field[:, 7] = deposit_species7(factor, sim)
    return field.reshape(-1), occ, sim


# ---------------------------------------------------------------------------
# Cell workloads

def seed_positions(dims, seed, band, count, margin, z_lo, z_hi, cell_at=None, index_base=0):
    """seed_cells placement (crates/marl-sim/src/seeding.rs:28-58) through the C ABI."""
    import ctypes

    from . import _abi

    lib = _abi.load_library()
    nv = dims[0] * dims[1] * dims[2]
    ca = np.full(nv, -1, np.int32) if cell_at is None else cell_at
    pos = np.zeros((max(count, 1), 3), np.uint16)
    lin = np.zeros(max(count, 1), np.uint64)
    placed = ctypes.c_int64()
    rc = lib.marl_seed_positions(ctypes.byref(_abi.marl_dims(*dims)), seed, band, count, margin, z_lo, z_hi,
                                 _abi.ptr(ca), _abi.ptr(pos), _abi.ptr(lin), index_base, ctypes.byref(placed))
    _abi.check(rc, what="marl_seed_positions")
    return pos[: placed.value], lin[: placed.value], ca


def random_reactions(rules, rng, vmax_range, km_range, light_p=0.0, k_cat=0.1, groups=None):
    """Overwrite the 16 reactions of each ruleset with random ones.

    The workload's "stoichiometry +-1 over 1-3 species" maps onto the reference
    Reaction (cell.rs:84-92) as: 2 species -> {substrate, product}; 3 species ->
    {substrate, product, cofactor}.  One species has no encoding (substrate ==
    product is skipped, cell.rs:250), so the count is drawn from {2, 3}.  The
    catalyst is species 15 (light) with probability light_p, else uniform over
    0..14; k_cat = 0.1; v_max and k_m are log-uniform over the given ranges.

    groups: optional int array (one entry per ruleset).  When given, the wiring
    (substrate, product, cofactor, catalyst) is drawn once per group and shared
    by its rulesets, as in a population grown from a few starter metabolisms;
    v_max and k_m stay per ruleset.
    """
    n = rules.shape[0]
    R = 16
    lv = np.log(vmax_range)
    lk = np.log(km_range)
    rx = rules["reactions"]
    vmax = np.exp(rng.uniform(lv[0], lv[1], (n, R))).astype(F32)
    km = np.exp(rng.uniform(lk[0], lk[1], (n, R))).astype(F32)
    m = n if groups is None else int(np.max(groups)) + 1
    nspec = rng.integers(2, 4, (m, R))
    sub = rng.integers(0, 15, (m, R))
    prod = (sub + rng.integers(1, 15, (m, R))) % 15
    cof = (sub + rng.integers(1, 15, (m, R))) % 15
    cof = np.where(cof == prod, (cof + 1) % 15, cof)
    cof = np.where(cof == sub, (cof + 1) % 15, cof)
    cof = np.where(nspec == 3, cof, 0xFF)
    cat = np.where(rng.random((m, R)) < light_p, 15, rng.integers(0, 15, (m, R)))
    if groups is not None:
        g = np.asarray(groups)
        sub, prod, cof, cat = sub[g], prod[g], cof[g], cat[g]
    rx["substrate"] = sub
    rx["product"] = prod
    rx["catalyst"] = cat
    rx["cofactor"] = cof
    rx["k_m"] = km
    rx["v_max"] = vmax
    rx["k_cat"] = k_cat


def linear_gradient_field(dims):
    """Species 1 and 3 fall linearly from 1 at z=0 to 0 at z=Z-1, species 2 rises
    from 0 to 1; the other nine are 0."""
    X, Y, Z = dims
    z = np.arange(Z, dtype=np.float64) / max(Z - 1, 1)
    f = np.zeros((Z, Y, X, 12), F32)
    f[..., 1] = (1.0 - z)[:, None, None]
    f[..., 3] = (1.0 - z)[:, None, None]
    f[..., 2] = z[:, None, None]
    return f.reshape(-1)


def winogradsky_column(dims=(128, 128, 64), seed=0, cells_per_band=1366,
                       bands=((4, 12), (24, 40), (52, 60)), starter_state=False):
    """The default-size column with about 4,096 cells.

    Linear-gradient field; cells seeded by seed_cells in three depth bands, one
    metabolism per band (phototroph, chemolithotroph, anaerobe, in that order)
    with that starter's receptors, transport, effectors and fate; 16 internals
    uniform in [0, 0.1]; 16 random reactions (v_max log-U[1e-3, 1e-1], k_m
    log-U[1e-2, 1], catalyst uniform).  Reference defaults otherwise
    (10 substeps).

    starter_state: keep each starter's own initial internals and reactions
    (starter_metabolisms.rs) instead of the random ones — the reference's own
    default population, which lives and divides for many ticks.
    """
    from .config import SimulationConfig
    from .starters import make_cell, make_ruleset
    from ._abi import CELL_DTYPE, RULESET_DTYPE

    sim = SimulationConfig(rng_seed=seed)
    rng = np.random.Generator(np.random.Philox(seed + 1))
    nv = dims[0] * dims[1] * dims[2]
    ca = np.full(nv, -1, np.int32)
    cells, rules = [], []
    for kind, (lo, hi) in enumerate(bands):
        pos, lin, ca = seed_positions(dims, seed, kind, cells_per_band, sim.seed_margin, lo, hi, ca,
                                      index_base=sum(len(c) for c in cells))
        c = np.zeros(len(pos), CELL_DTYPE)
        r = np.zeros(len(pos), RULESET_DTYPE)
        r[:] = make_ruleset(kind)
        for k in range(len(pos)):
            c[k] = make_cell(kind, pos[k], int(lin[k]))
        cells.append(c)
        rules.append(r)
    cells = np.concatenate(cells)
    rules = np.concatenate(rules)
    if not starter_state:
        cells["internal"] = rng.uniform(0.0, 0.1, (len(cells), 16)).astype(F32)
        random_reactions(rules, rng, (1e-3, 1e-1), (1e-2, 1.0))
    return linear_gradient_field(dims), cells, rules, sim


def kinetics_at_scale(dims=(256, 256, 128), n_cells=1_000_000, seed=0, shared_wiring=False):
    """Plain per-cell kinetics at scale.

    n_cells placed uniformly (seed_cells with margin 0 over all z); starter type
    uniform over the three metabolisms (their transport, effectors and fate
    thresholds, i.e. "as in the reference"); 16 internals uniform in [0, 1];
    16 random reactions (v_max log-U[1e-3, 1], k_m log-U[1e-3, 1], 2-3 species,
    light-catalysed with p = 0.2); field species uniform in [0, 1] with species 7
    in [0, 0.1].  Reference defaults otherwise (10 substeps, mutation_rate 0.05).
    shared_wiring: the reaction wiring is drawn once per metabolism (three
    reaction topologies) instead of once per cell — the surrogate's workload.
    """
    from .config import SimulationConfig
    from .starters import make_ruleset
    from ._abi import CELL_DTYPE, RULESET_DTYPE

    sim = SimulationConfig(rng_seed=seed)
    rng = np.random.Generator(np.random.Philox(seed + 2))
    pos, lin, _ = seed_positions(dims, seed, 7, n_cells, 0, 0, dims[2])
    n = len(pos)
    kinds = rng.integers(0, 3, n)
    cells = np.zeros(n, CELL_DTYPE)
    rules = np.zeros(n, RULESET_DTYPE)
    for kind in range(3):
        sel = kinds == kind
        rules[sel] = make_ruleset(kind)
    cells["pos"] = pos
    cells["lineage_id"] = lin
    cells["starter_type"] = kinds
    cells["internal"] = rng.random((n, 16), dtype=F32)
    random_reactions(rules, rng, (1e-3, 1.0), (1e-3, 1.0), light_p=0.2, groups=kinds if shared_wiring else None)
    nv = dims[0] * dims[1] * dims[2]
    field = rng.random((nv, 12), dtype=F32)
    field[:, 7] *= 0.1
    return field.reshape(-1), cells, rules, sim


def topology_hash(rules) -> np.ndarray:
    """64-bit FNV-1a over each ruleset's wiring: per reaction the word
    substrate | product << 8 | catalyst << 16 | cofactor << 24, then per
    transporter ext | int << 8, then per effector int | ext << 8 — the key the
    surrogate's topology classes use."""
    r = np.atleast_1d(rules)
    rx = r["reactions"]
    words = [rx["substrate"].astype(np.uint64) | rx["product"].astype(np.uint64) << np.uint64(8)
             | rx["catalyst"].astype(np.uint64) << np.uint64(16) | rx["cofactor"].astype(np.uint64) << np.uint64(24)]
    tp, ef = r["transport"], r["effectors"]
    words.append(tp["ext_species"].astype(np.uint64) | tp["int_species"].astype(np.uint64) << np.uint64(8))
    words.append(ef["int_species"].astype(np.uint64) | ef["ext_species"].astype(np.uint64) << np.uint64(8))
    w = np.concatenate(words, axis=1)
    h = np.full(len(r), 1469598103934665603, np.uint64)
    with np.errstate(over="ignore"):
        for k in range(w.shape[1]):
            h = (h ^ w[:, k]) * np.uint64(1099511628211)
    return h


def ensemble_columns(n_columns=32, seed=0, sizes=(64, 96, 128, 192), cells_range=(1000, 200000), max_fill=0.15):
    """Independent columns of uneven sizes and seeds.

    Column c: x = y drawn from `sizes`, z = x / 2; its own rng seed
    (seed * 1000 + c); a cell count log-uniform over `cells_range`, capped at
    `max_fill` of the voxels; cells placed uniformly, each one of the three
    starter metabolisms with that starter's own initial state; linear-gradient
    field.  Returns [(dims, field, cells, rules, rng_seed)]."""
    from .starters import make_cells

    out = []
    for c in range(n_columns):
        rng = np.random.Generator(np.random.Philox(seed * 1000 + c + 17))
        x = int(rng.choice(sizes))
        dims = (x, x, x // 2)
        nv = dims[0] * dims[1] * dims[2]
        lo, hi = np.log(cells_range[0]), np.log(cells_range[1])
        n = min(int(np.exp(rng.uniform(lo, hi))), int(max_fill * nv))
        vox = np.sort(rng.choice(nv, n, replace=False))
        pos = np.stack([vox % dims[0], (vox // dims[0]) % dims[1], vox // (dims[0] * dims[1])], 1)
        kinds = rng.integers(0, 3, n)
        cells, rules = make_cells(kinds, pos, rng.integers(0, 2**63, n, dtype=np.uint64))
        out.append((dims, linear_gradient_field(dims), cells, rules, seed * 1000 + c))
    return out
