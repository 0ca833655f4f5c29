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
