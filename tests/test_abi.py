"""The C-ABI library loads, exports every declared symbol, and its record
layouts agree with include/marl_hip.h (CPU only: no compute calls)."""
import ctypes
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "marl_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(marl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load_library()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(pkg._abi.SIGNATURES), set(names) ^ set(pkg._abi.SIGNATURES)


def test_record_layouts_match_header(pkg, tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "marl_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(marl_sim_params),"
        " sizeof(marl_cell_state), sizeof(marl_ruleset), offsetof(marl_ruleset, division_energy),"
        " offsetof(marl_cell_state, lineage_id), sizeof(marl_tick_stats), offsetof(marl_sim_params, rng_seed),"
        " sizeof(marl_train_config), sizeof(marl_train_report), sizeof(marl_surr_stats),"
        " offsetof(marl_surr_stats, err_hist)); return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    A = pkg._abi
    want = [ctypes.sizeof(A.marl_sim_params), A.CELL_DTYPE.itemsize, A.RULESET_DTYPE.itemsize,
            A.RULESET_DTYPE.fields["division_energy"][1], A.CELL_DTYPE.fields["lineage_id"][1],
            ctypes.sizeof(A.marl_tick_stats), A.marl_sim_params.rng_seed.offset,
            ctypes.sizeof(A.marl_train_config), ctypes.sizeof(A.marl_train_report), ctypes.sizeof(A.marl_surr_stats),
            A.marl_surr_stats.err_hist.offset]
    assert got == want


def test_no_device_maps_to_no_adapter(pkg):
    # ≙ the reference's skip path when wgpu finds no adapter (gpu_diffusion.rs:99-106)
    lib = pkg.load_library()
    if lib.marl_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(pkg.NoAdapter):
        pkg.GpuFieldDiffuser.new()


def test_invalid_dims_rejected_without_device(pkg):
    lib = pkg.load_library()
    h = ctypes.c_void_p()
    for dims in [(100, 8, 8), (0, 8, 8), (64, -1, 8)]:
        d = pkg._abi.marl_dims(*dims)
        assert lib.marl_create(ctypes.byref(d), ctypes.byref(h)) == pkg._abi.ERR_INVALID_INPUT


def test_sim_params_defaults_match_reference(pkg):
    # crates/marl-config/src/lib.rs:108-168
    p = pkg.SimulationConfig().to_params()
    assert p.diffusion_substeps == 10 and p.dt == 1.0
    assert list(np.float32(p.d_voxel)) == list(np.float32([0.0, 1.5, 1.0, 1.2, 0.8, 0.5, 0.5, 0.1, 0.3, 0.3, 0.3, 0.3]))
    assert np.float32(p.alpha_eps) == np.float32(0.8) and np.float32(p.k_eps) == np.float32(2.0)
    assert np.float32(p.light_floor) == np.float32(1e-7) and p.division_neighbor_distance == 2
