"""Pins the oracle's stencil, boundary sources and light scan (CPU only).

Checks come from the reference's own tests (field.rs:270-312, the six cases of
gpu_diffusion.rs:114-150) and from what the mathematics fixes: a uniform
field has a zero Laplacian, an impulse spreads with weights dt*D to its six
faces, walls and occupied faces are Neumann, mirror symmetry, FP64 mass
conservation with a uniform niche factor, and the Beer-Lambert closed form.
"""
import numpy as np
import pytest

from fixtures import REFERENCE_CASES, init_deterministic, init_wave, occupancy_center, occupancy_empty, vidx

DIMS = (128, 128, 64)  # the reference's compile-time grid (marl-config/src/lib.rs:17-19)
SMALL = (32, 16, 12)


def test_reference_occupied_voxel_is_copied_unchanged(pkg, orc):
    # crates/marl-field/src/field.rs:270-289
    sim = pkg.SimulationConfig(diffusion_substeps=1)
    f = init_deterministic(DIMS)
    occ = occupancy_center(DIMS)
    out = orc.diffuse_tick(f, occ, DIMS, sim)
    v = vidx(DIMS, 64, 64, 32)
    assert np.array_equal(out[v * 12:(v + 1) * 12], f[v * 12:(v + 1) * 12])


def test_reference_deterministic_diffusion_stays_finite_and_nonnegative(pkg, orc):
    # crates/marl-field/src/field.rs:291-312
    sim = pkg.SimulationConfig(diffusion_substeps=1)
    out = orc.diffuse_tick(init_deterministic(DIMS), occupancy_empty(DIMS), DIMS, sim)
    assert np.isfinite(out).all() and (out >= 0).all()


@pytest.mark.parametrize("name,occ_fn,substeps", REFERENCE_CASES)
def test_reference_cases_f32_within_reference_tolerance_of_f64(pkg, orc, name, occ_fn, substeps):
    # the reference's agreement bar is max abs <= 1e-4 (gpu_diffusion.rs:111); FP32 vs FP64 here
    sim = pkg.SimulationConfig(diffusion_substeps=substeps)
    f = init_wave(DIMS)
    occ = occ_fn(DIMS)
    f32 = orc.diffuse_tick(f, occ, DIMS, sim)
    f64 = f.astype(np.float64)
    dt_sub = np.float64(np.float32(sim.dt) / np.float32(substeps))
    for _ in range(substeps):
        f64 = orc.substep_f64(f64, occ, DIMS, sim, dt_sub)
    assert np.abs(f32 - f64).max() <= 1e-4


def test_uniform_field_follows_pure_decay(pkg, orc):
    # a uniform field has a zero Laplacian: c' = c * (1 - dt_sub * lambda_s)
    sim = pkg.SimulationConfig(diffusion_substeps=1)
    nv = np.prod(SMALL)
    f = np.full(nv * 12, 0.5, np.float32)
    out = orc.diffuse_tick(f, None, SMALL, sim).reshape(nv, 12)
    want = 0.5 * (1.0 - np.asarray(sim.lambda_decay, np.float64))
    assert np.allclose(out, want[None, :], rtol=1e-6, atol=0)


@pytest.mark.parametrize("where", ["interior", "x0_wall", "z_top_wall", "next_to_occupied"])
def test_impulse_weights(pkg, orc, where):
    # one unit of species 2 in an otherwise empty field (species 7 = 0 -> niche factor 1)
    X, Y, Z = SMALL
    sim = pkg.SimulationConfig(dt=0.1, diffusion_substeps=1, lambda_decay=[0.0] * 12)
    dt, d = 0.1, sim.d_voxel[2]
    cx, cy, cz = {"interior": (10, 7, 5), "x0_wall": (0, 7, 5), "z_top_wall": (10, 7, 0),
                  "next_to_occupied": (10, 7, 5)}[where]
    nv = X * Y * Z
    f = np.zeros(nv * 12, np.float32)
    f[vidx(SMALL, cx, cy, cz) * 12 + 2] = 1.0
    occ = np.zeros(nv, np.uint8)
    if where == "next_to_occupied":
        occ[vidx(SMALL, cx, cy + 1, cz)] = 1
    out = orc.diffuse_tick(f, occ, SMALL, sim).reshape(Z, Y, X, 12)[..., 2]
    open_faces = 6 - (where != "interior")
    assert out[cz, cy, cx] == pytest.approx(1.0 - dt * d * open_faces, rel=1e-6)
    nbrs = [(cx + 1, cy, cz), (cx - 1, cy, cz), (cx, cy + 1, cz), (cx, cy - 1, cz), (cx, cy, cz + 1), (cx, cy, cz - 1)]
    total = out[cz, cy, cx]
    for (x, y, z) in nbrs:
        if 0 <= x < X and 0 <= y < Y and 0 <= z < Z:
            if occ[vidx(SMALL, x, y, z)]:
                assert out[z, y, x] == 0.0
            else:
                assert out[z, y, x] == pytest.approx(dt * d, rel=1e-6)
                total += out[z, y, x]
    assert total == pytest.approx(1.0, rel=1e-6)  # nothing leaves through walls or cells
    assert np.count_nonzero(out) == 1 + open_faces


def test_mirror_symmetry_is_bitwise(pkg, orc):
    X, Y, Z = SMALL
    rng = np.random.default_rng(3)
    half = rng.random((Z, Y, X // 2, 12), dtype=np.float32)
    f = np.concatenate([half, half[:, :, ::-1]], axis=2)
    oh = (rng.random((Z, Y, X // 2)) < 0.1).astype(np.uint8)
    occ = np.concatenate([oh, oh[:, :, ::-1]], axis=2)
    sim = pkg.SimulationConfig(diffusion_substeps=3)
    out = orc.diffuse_tick(f.reshape(-1), occ.reshape(-1), SMALL, sim).reshape(Z, Y, X, 12)
    assert np.array_equal(out, out[:, :, ::-1])


@pytest.mark.parametrize("with_cells", [False, True])
def test_f64_mass_is_conserved_with_uniform_niche_and_no_decay(pkg, orc, with_cells):
    X, Y, Z = SMALL
    nv = X * Y * Z
    rng = np.random.default_rng(5)
    f = rng.random((nv, 12)) * 0.5 + 0.25
    f[:, 7] = 0.7
    sim = pkg.SimulationConfig(diffusion_substeps=1, lambda_decay=[0.0] * 12)
    occ = (rng.random(nv) < 0.15).astype(np.uint8) if with_cells else None
    out = orc.substep_f64(f.reshape(-1), occ, SMALL, sim, 0.1).reshape(nv, 12)
    assert np.allclose(out.sum(0), f.sum(0), rtol=1e-12)


def test_f64_mass_not_conserved_when_niche_varies(pkg, orc):
    # crates/marl-field/src/field.rs:141-154 scales the flux by the centre voxel's factor only
    X, Y, Z = SMALL
    nv = X * Y * Z
    rng = np.random.default_rng(6)
    f = rng.random((nv, 12)) + 0.5
    f[:, 7] = rng.random(nv) * 4.0
    sim = pkg.SimulationConfig(diffusion_substeps=1, lambda_decay=[0.0] * 12)
    out = orc.substep_f64(f.reshape(-1), None, SMALL, sim, 0.1).reshape(nv, 12)
    assert abs(out[:, 1].sum() - f[:, 1].sum()) > 1e-9 * f[:, 1].sum()


def test_boundary_sources(pkg, orc):
    X, Y, Z = SMALL
    nv = X * Y * Z
    f = np.full((nv, 12), 9.9, np.float32)
    f[:, 3] = 9.0
    sim = pkg.SimulationConfig()
    out = orc.boundary(f.reshape(-1), SMALL, sim).reshape(Z, Y, X, 12)
    assert np.all(out[0, :, :, 1] == np.float32(10.0))   # min(9.9 + 0.4, c_max)
    assert np.all(out[0, :, :, 3] == np.float32(9.0) + np.float32(0.15))
    assert np.all(out[Z - 1, :, :, 2] == np.float32(10.0))  # min(9.9 + 0.5, c_max)
    rest = out.copy()
    rest[0, :, :, 1] = rest[Z - 1, :, :, 2] = 9.9
    rest[0, :, :, 3] = 9.0
    assert np.array_equal(rest.reshape(nv, 12), f)


def test_light_matches_beer_lambert_closed_form(pkg, orc):
    X, Y, Z = SMALL
    nv = X * Y * Z
    a = 0.37
    f = np.zeros((nv, 12), np.float32)
    f[:, 4] = a
    occ = np.zeros((Z, Y, X), np.uint8)
    occ[3, 2, 5] = occ[7, 2, 5] = 1
    sim = pkg.SimulationConfig(chemical_absorption=0.05, cell_absorption=0.3, light_floor=1e-7)
    lt = orc.light(f.reshape(-1), occ.reshape(-1), SMALL, sim).reshape(Z, Y, X)
    z = np.arange(Z)
    want = np.exp(-0.05 * np.float64(np.float32(a)) * z)
    assert np.allclose(lt[:, 0, 0], want, rtol=2e-6)
    cells_above = (z > 3).astype(int) + (z > 7).astype(int)
    assert np.allclose(lt[:, 2, 5], want * np.exp(-0.3 * cells_above), rtol=2e-6)
    lt64 = orc.light_f64(f.reshape(-1).astype(np.float64), occ.reshape(-1), SMALL, sim)
    assert np.allclose(lt.reshape(-1), lt64, rtol=2e-6)


def test_light_floor_switches_to_zero(pkg, orc):
    X, Y, Z = SMALL
    nv = X * Y * Z
    f = np.zeros((nv, 12), np.float32)
    f[:, 4] = 10.0
    sim = pkg.SimulationConfig(chemical_absorption=0.5, light_floor=1e-7)  # e^-5 per voxel
    lt = orc.light(f.reshape(-1), None, SMALL, sim).reshape(Z, Y, X)
    col = lt[:, 0, 0].astype(np.float64)
    first_zero = int(np.argmax(col == 0))
    assert first_zero == 4  # e^-20 = 2.1e-9 < 1e-7 after four voxels, then it stays 0
    assert np.all(col[first_zero:] == 0) and np.all(col[:first_zero] > 1e-7)
