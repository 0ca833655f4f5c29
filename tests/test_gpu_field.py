"""GPU parity of the field path through the C ABI: masked stencil, boundary
sources, light.  The stencil and light kernels compute in the reference's f32
operator order without contraction, so the bar is bit-identity with the
oracle (stronger than the reference's own 1e-4 CPU/GPU tolerance,
crates/marl-gpu/tests/gpu_diffusion.rs:111)."""
import numpy as np
import pytest

from fixtures import REFERENCE_CASES, init_deterministic, init_wave, occupancy_center, occupancy_empty, vidx

pytestmark = pytest.mark.gpu

DIMS = (128, 128, 64)


def _rand_case(dims, seed, occ_frac=0.1):
    rng = np.random.default_rng(seed)
    nv = dims[0] * dims[1] * dims[2]
    f = rng.random(nv * 12, dtype=np.float32)
    f.reshape(nv, 12)[:, 7] *= 3.0
    occ = (rng.random(nv) < occ_frac).astype(np.uint8)
    return f, occ


@pytest.mark.parametrize("name,occ_fn,substeps", REFERENCE_CASES)
def test_reference_cases_drop_in(pkg, orc, name, occ_fn, substeps):
    """gpu_diffusion.rs:114-150 through the drop-in entry point."""
    sim = pkg.SimulationConfig(diffusion_substeps=substeps)
    f = init_wave(DIMS)
    occ = occ_fn(DIMS)
    want = orc.diffuse_tick(f, occ, DIMS, sim)
    d = pkg.GpuFieldDiffuser.new(DIMS)
    got = f.copy()
    d.diffuse_tick_with_cells(got, occ, sim)
    d.close()
    assert np.abs(got - want).max() <= 1e-4, name
    assert np.array_equal(got, want), f"{name}: {np.count_nonzero(got != want)} values differ"


def test_reference_field_tests_on_gpu(pkg):
    """field.rs:270-312 on the GPU path."""
    sim = pkg.SimulationConfig(diffusion_substeps=1)
    d = pkg.GpuFieldDiffuser.new(DIMS)
    f = init_deterministic(DIMS)
    before = f.copy()
    d.diffuse_tick_with_cells(f, occupancy_center(DIMS), sim)
    v = vidx(DIMS, 64, 64, 32)
    assert np.array_equal(f[v * 12:(v + 1) * 12], before[v * 12:(v + 1) * 12])
    f = init_deterministic(DIMS)
    d.diffuse_tick_with_cells(f, occupancy_empty(DIMS), sim)
    assert np.isfinite(f).all() and (f >= 0).all()
    d.close()


def test_drop_in_contract(pkg):
    """Length checks (field_diffusion.rs:164-175), substeps==0 no-op (:176-178),
    host buffer untouched on error."""
    d = pkg.GpuFieldDiffuser.new(DIMS)
    f = init_wave(DIMS)
    keep = f.copy()
    with pytest.raises(pkg.InvalidInput):
        d.diffuse_tick_with_cells(f[:-12].copy(), occupancy_empty(DIMS), pkg.SimulationConfig())
    with pytest.raises(pkg.InvalidInput):
        d.diffuse_tick_with_cells(f, occupancy_empty(DIMS)[:-1], pkg.SimulationConfig())
    assert np.array_equal(f, keep)
    d.diffuse_tick_with_cells(f, occupancy_empty(DIMS), pkg.SimulationConfig(diffusion_substeps=0))
    assert np.array_equal(f, keep)
    d.close()


@pytest.mark.parametrize("dims", [(96, 70, 37), (64, 64, 16), (32, 5, 3)])
def test_every_tile_shape_bit_identical(pkg, orc, dims):
    """Partial x/y tiles (X=96, Y=70), partial z-chunks (Z=37), tiny grids."""
    f, occ = _rand_case(dims, 1)
    sim = pkg.SimulationConfig(diffusion_substeps=3)
    want = orc.diffuse_tick(f, occ, dims, sim)
    with pkg.MarlGpu(dims) as g:
        n = g.lib.marl_stencil_tile_count()
        for t in range(n):
            g.upload_field_aos(f)
            g.upload_occupancy(occ)
            g.tile = t
            g.diffuse(sim)
            got = g.download_field_aos()
            assert np.array_equal(got, want), f"tile {t} {g.tile_shape(t)}: {np.count_nonzero(got != want)} differ"


def test_soa_roundtrip_and_occupancy_roundtrip(pkg):
    dims = (64, 32, 8)
    f, occ = _rand_case(dims, 2)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(f)
        soa = g.download_field_soa()
        assert np.array_equal(soa.reshape(12, -1).T.reshape(-1), f)
        g.upload_field_soa(soa)
        assert np.array_equal(g.download_field_aos(), f)
        g.upload_occupancy(occ)
        assert np.array_equal(g.download_occupancy(), occ)


def test_boundary_sources_bit_identical(pkg, orc):
    dims = (64, 48, 20)
    f, _ = _rand_case(dims, 3)
    f *= 10.0
    sim = pkg.SimulationConfig()
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(f)
        g.apply_boundary_sources(sim)
        assert np.array_equal(g.download_field_aos(), orc.boundary(f, dims, sim))


def test_light_bit_identical(pkg, orc):
    dims = (96, 40, 50)
    f, occ = _rand_case(dims, 4, occ_frac=0.2)
    f.reshape(-1, 12)[:, 4] *= 4.0
    for sim in [pkg.SimulationConfig(), pkg.SimulationConfig(chemical_absorption=0.4, cell_absorption=0.5)]:
        want = orc.light(f, occ, dims, sim)
        with pkg.MarlGpu(dims) as g:
            g.upload_field_aos(f)
            g.upload_occupancy(occ)
            g.light_update(sim)
            got = g.download_light()
        assert np.array_equal(got, want), np.count_nonzero(got != want)
        want64 = orc.light_f64(f.astype(np.float64), occ, dims, sim)
        assert np.allclose(got, want64, rtol=5e-6, atol=1e-30)


def test_full_size_stress_grid_two_substeps(pkg, orc):
    """256x256x128 (the stencil stress grid) with 5% occupancy: bit-identical."""
    dims = (256, 256, 128)
    f, occ = _rand_case(dims, 5, occ_frac=0.05)
    sim = pkg.SimulationConfig(diffusion_substeps=8, d_voxel=[0.05 + 0.01 * s for s in range(12)])
    want = orc.substep(orc.substep(f, occ, dims, sim, np.float32(0.125)), occ, dims, sim, np.float32(0.125))
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(f)
        g.upload_occupancy(occ)
        g.diffuse(sim, 2)
        got = g.download_field_aos()
    assert np.array_equal(got, want), np.count_nonzero(got != want)


def test_autotune_keeps_field_and_picks_valid_tile(pkg, orc):
    dims = (128, 64, 32)
    f, occ = _rand_case(dims, 6)
    sim = pkg.SimulationConfig(diffusion_substeps=2)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(f)
        g.upload_occupancy(occ)
        t = g.autotune(sim)
        assert 0 <= t < g.lib.marl_stencil_tile_count()
        assert np.array_equal(g.download_field_aos(), f)
        g.diffuse(sim)
        assert np.array_equal(g.download_field_aos(), orc.diffuse_tick(f, occ, dims, sim))


@pytest.mark.parametrize("dims", [(128, 70, 37), (64, 64, 16), (32, 5, 3)])
def test_temporal_tiles_bit_identical(pkg, orc, dims):
    """Two-substep tiles: even and odd substep counts (an odd remainder runs a
    single-substep launch), and a full tick with the boundary sources fused into
    the first launch, against the oracle."""
    f, occ = _rand_case(dims, 7)
    with pkg.MarlGpu(dims) as g:
        temporal = [t for t in range(g.lib.marl_stencil_tile_count()) if g.tile_substeps(t) == 2]
        assert temporal
        for n in (1, 2, 4, 5):
            sim = pkg.SimulationConfig(diffusion_substeps=n)
            want = orc.diffuse_tick(f, occ, dims, sim)
            for t in temporal:
                g.upload_field_aos(f)
                g.upload_occupancy(occ)
                g.tile = t
                g.diffuse(sim)
                got = g.download_field_aos()
                assert np.array_equal(got, want), f"tile {t}, {n} substeps: {np.count_nonzero(got != want)} differ"
        A = pkg._abi
        empty_c = np.zeros(0, A.CELL_DTYPE)
        empty_r = np.zeros(0, A.RULESET_DTYPE)
        sim = pkg.SimulationConfig(diffusion_substeps=4)
        fw, lw, _, _, _, _ = orc.tick(dims, sim, 0, f, empty_c, empty_r)
        for t in temporal:
            g.upload_field_aos(f)
            g.upload_cells(empty_c, empty_r)
            g.tile = t
            g.tick(sim, 0, 1)
            assert np.array_equal(g.download_field_aos(), fw), t
            assert np.array_equal(g.download_light(), lw), t
            g.upload_field_aos(f)
            g.tick_graph(sim, 0, 1)
            assert np.array_equal(g.download_field_aos(), fw), t


def test_calibration_probes_report_their_bytes(pkg):
    """The copy probe moves the field's bytes twice per rep; the gather probe
    reads 2^24 / 4 groups x 4 sectors x 64 B per rep; both take measurable time
    and leave the resident field as it was."""
    import ctypes

    dims = (128, 128, 64)
    field = np.random.default_rng(3).random(np.prod(dims) * 12).astype(np.float32)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        ms, by = ctypes.c_float(), ctypes.c_double()
        g._chk(g.lib.marl_copy_probe(g.h, 3, ctypes.byref(ms), ctypes.byref(by)))
        assert ms.value > 0 and by.value == 2.0 * field.size * 4
        g._chk(g.lib.marl_gather_probe(g.h, 3, ctypes.byref(ms), ctypes.byref(by)))
        assert ms.value > 0 and by.value == (1 << 24) // 4 * 4 * 64.0
        assert np.array_equal(g.download_field_aos(), field)


def test_mfma_probe_reports_a_plausible_rate(pkg):
    """The dense BF16 calibration: FLOP per rep = 4 blocks per CU x 4 waves x
    4,096 steps x 4 chains x 32,768 FLOP, and a rate between a tenth of the
    vendor's 2.5 PFLOP/s and that peak plus a margin."""
    import ctypes
    with pkg.MarlGpu((64, 64, 32)) as g:
        ms, fl = ctypes.c_float(), ctypes.c_double()
        g._chk(g.lib.marl_mfma_probe(g.h, 2, ctypes.byref(ms), ctypes.byref(fl)))
        per_cu = 4 * 4 * 4096 * 4 * 32768.0
        cus = fl.value / per_cu
        assert cus == int(cus) and cus >= 1
        tflops = 2 * fl.value / (ms.value * 1e-3) / 1e12
        print("mfma probe", round(tflops, 1), "TFLOP/s on", int(cus), "CUs", flush=True)
        assert 250.0 < tflops < 3000.0
