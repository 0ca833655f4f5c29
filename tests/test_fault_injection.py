"""Failure handling through the C ABI's fault-injection hook (MARL_HIP_FAULT).

The reference engine falls back to CPU diffusion when GPU init fails and,
when a per-tick GPU call fails, redoes that tick's diffusion on the CPU for the
rest of the run (crates/marl-sim/src/lib.rs:133-146, 180-191); the fallback is
exact because the host field changes only after a successful readback
(crates/marl-gpu/src/field_diffusion.rs:285-290).  These tests drive that
protocol from the caller's side, with the oracle's CPU diffusion standing in
for the reference's CPU path."""
import ctypes

import numpy as np
import pytest

F = np.float32


def test_injected_init_fault_is_device_request(pkg, monkeypatch):
    """MARL_HIP_FAULT=init fails marl_create with DEVICE_INIT (≙ GpuError::DeviceRequest)
    and a message, with or without a GPU."""
    monkeypatch.setenv("MARL_HIP_FAULT", "init")
    lib = pkg.load_library()
    h = ctypes.c_void_p()
    d = pkg._abi.marl_dims(64, 8, 8)
    assert lib.marl_create(ctypes.byref(d), ctypes.byref(h)) == pkg._abi.ERR_DEVICE_INIT
    assert b"injected fault" in lib.marl_last_error(None)
    with pytest.raises(pkg.DeviceRequest, match="injected fault"):
        pkg.GpuFieldDiffuser.new((64, 8, 8))


def _inputs(dims, seed):
    rng = np.random.default_rng(seed)
    nv = dims[0] * dims[1] * dims[2]
    return rng.random(nv * 12, dtype=F), (rng.random(nv) < 0.1).astype(np.uint8)


@pytest.mark.gpu
def test_failed_dropin_tick_leaves_field_and_cpu_fallback_is_exact(pkg, orc, monkeypatch):
    dims = (64, 64, 32)
    field, occ = _inputs(dims, 3)
    sim = pkg.SimulationConfig(diffusion_substeps=4)
    want = field.copy()
    for _ in range(5):
        want = orc.diffuse_tick(want, occ, dims, sim)
    monkeypatch.setenv("MARL_HIP_FAULT", "tick:3")
    f = field.copy()
    gpu = pkg.GpuFieldDiffuser.new(dims)
    failed_at = None
    try:
        for t in range(5):
            if gpu is not None:
                before = f.copy()
                try:
                    gpu.diffuse_tick_with_cells(f, occ, sim)
                    continue
                except pkg.BufferMap as e:
                    assert "injected fault" in str(e)
                    assert np.array_equal(f.view(np.uint32), before.view(np.uint32))  # host field untouched
                    failed_at = t
                    gpu.close()
                    gpu = None
            f = orc.diffuse_tick(f, occ, dims, sim)   # the CPU path for this and later ticks
    finally:
        if gpu is not None:
            gpu.close()
    assert failed_at == 2
    assert np.array_equal(f.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_failed_resident_tick_changes_no_state(pkg, monkeypatch):
    from test_gpu_cells import assert_cells_equal, population

    A = pkg._abi
    dims = (64, 32, 16)
    cells, rules = population(A, dims, 2000, 8)
    field, _ = _inputs(dims, 9)
    sim = pkg.SimulationConfig(diffusion_substeps=2)
    with pkg.MarlGpu(dims) as g, pkg.MarlGpu(dims) as ref:
        for x in (g, ref):
            x.upload_field_aos(field)
            x.upload_cells(cells, rules)
        monkeypatch.setenv("MARL_HIP_FAULT", "tick:2")
        g.tick(sim, 0, 1)
        f1, (c1, r1) = g.download_field_aos(), g.download_cells()
        with pytest.raises(pkg.BufferMap, match="injected fault"):
            g.tick(sim, 1, 1)
        assert np.array_equal(g.download_field_aos().view(np.uint32), f1.view(np.uint32))
        c2, r2 = g.download_cells()
        assert_cells_equal(c2, r2, c1, r1)
        g.tick(sim, 1, 1)          # the retry is the tick the failed call would have run
        monkeypatch.delenv("MARL_HIP_FAULT")
        ref.tick(sim, 0, 2)
        ca, ra = ref.download_cells()
        cb, rb = g.download_cells()
        assert_cells_equal(cb, rb, ca, ra)
        assert np.array_equal(g.download_field_aos().view(np.uint32), ref.download_field_aos().view(np.uint32))
