"""SimulationConfig: the runtime parameters the tick path reads.

Field names and defaults follow crates/marl-config/src/lib.rs:41-168 (and the
sample marl.toml).  `rng_seed` is new: the reference draws from an OS-seeded
ThreadRng (crates/marl-sim/src/lib.rs:35); this project keys a counter-based
Philox4x64-10 stream with it instead.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

from ._abi import S_EXT, marl_sim_params

# Compile-time constants of the reference (crates/marl-config/src/lib.rs:17-29);
# here the grid is chosen per context.
GRID_X, GRID_Y, GRID_Z = 128, 128, 64


@dataclass
class SimulationConfig:
    dx: float = 100.0e-6
    dt: float = 1.0
    diffusion_substeps: int = 10
    d_voxel: list = field(default_factory=lambda: [0.0, 1.5, 1.0, 1.2, 0.8, 0.5, 0.5, 0.1, 0.3, 0.3, 0.3, 0.3])
    lambda_decay: list = field(
        default_factory=lambda: [0.0, 0.01, 0.01, 0.005, 0.03, 0.05, 0.05, 0.002, 0.01, 0.01, 0.01, 0.01])
    source_rate_oxidant: float = 0.4
    source_rate_carbon: float = 0.15
    source_rate_reductant: float = 0.5
    epsilon: float = 0.001
    c_max: float = 10.0
    lambda_maintenance: float = 0.12
    hard_death_floor: float = 0.01
    reaction_maintenance: float = 0.003
    base_division_prep: float = 20.0
    prep_maintenance_multiplier: float = 2.0
    rush_penalty_rate: float = 0.05
    alpha_eps: float = 0.8
    k_eps: float = 2.0
    light_efficiency: float = 0.0
    surface_intensity: float = 1.0
    cell_absorption: float = 0.3
    chemical_absorption: float = 0.05
    light_floor: float = 1e-7
    mutation_stddev: float = 0.1
    structural_mutation_rate_mult: float = 0.1
    meta_mutation_rate: float = 0.01
    meta_mutation_clamp_low: float = 0.001
    meta_mutation_clamp_high: float = 0.5
    hill_exponent_clamp_low: float = 0.5
    hill_exponent_clamp_high: float = 8.0
    active_reaction_threshold: float = 1e-9
    seed_margin: int = 5
    phototroph_z_lo: float = 0.0
    phototroph_z_hi: float = 3.0
    chemolithotroph_z_lo: float = 80.0
    chemolithotroph_z_hi: float = 130.0
    anaerobe_z_lo: float = 120.0
    anaerobe_z_hi: float = 180.0
    division_neighbor_distance: int = 2
    boundary_prime_layers: int = 2
    boundary_prime_oxidant: float = 0.5
    boundary_prime_carbon: float = 0.3
    boundary_prime_reductant: float = 0.5
    rng_seed: int = 0

    def replace(self, **kw) -> "SimulationConfig":
        return dataclasses.replace(self, **kw)

    def to_params(self) -> marl_sim_params:
        """Pack into the C struct marl_sim_params (include/marl_hip.h)."""
        p = marl_sim_params()
        for f in marl_sim_params._fields_:
            name = f[0]
            if name.startswith("_"):
                continue
            v = getattr(self, name)
            if name in ("d_voxel", "lambda_decay"):
                if len(v) != S_EXT:
                    raise ValueError(f"{name} needs {S_EXT} values")
                arr = getattr(p, name)
                for i, x in enumerate(v):
                    arr[i] = float(x)
            else:
                setattr(p, name, v)
        return p
