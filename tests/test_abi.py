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
        # a failed create leaves its reason for marl_last_error(NULL)
        assert b"multiple of 32" in lib.marl_last_error(None)


def test_failed_create_keeps_its_message(pkg):
    lib = pkg.load_library()
    h = ctypes.c_void_p()
    d = pkg._abi.marl_dims(1024, 1024, 1024)   # a 48 GiB field buffer
    assert lib.marl_create(ctypes.byref(d), ctypes.byref(h)) == pkg._abi.ERR_INVALID_INPUT
    assert b"2 GiB" in lib.marl_last_error(None)
    if lib.marl_device_count() == 0:
        d = pkg._abi.marl_dims(64, 8, 8)
        assert lib.marl_create(ctypes.byref(d), ctypes.byref(h)) == pkg._abi.ERR_NO_DEVICE
        assert b"no GPU" in lib.marl_last_error(None)


def test_sim_params_defaults_match_reference(pkg):
    # crates/marl-config/src/lib.rs:108-168
    p = pkg.SimulationConfig().to_params()
    assert p.diffusion_substeps == 10 and p.dt == 1.0
    assert list(np.float32(p.d_voxel)) == list(np.float32([0.0, 1.5, 1.0, 1.2, 0.8, 0.5, 0.5, 0.1, 0.3, 0.3, 0.3, 0.3]))
    assert np.float32(p.alpha_eps) == np.float32(0.8) and np.float32(p.k_eps) == np.float32(2.0)
    assert np.float32(p.light_floor) == np.float32(1e-7) and p.division_neighbor_distance == 2


def test_entry_points_validate_arguments_without_device(pkg):
    """Argument checks that return before touching a device."""
    A = pkg._abi
    lib = pkg.load_library()
    out = ctypes.c_void_p()
    assert lib.marl_ensemble_create(None, 0, ctypes.byref(out)) == A.ERR_INVALID_INPUT
    assert lib.marl_ensemble_create(None, 3, ctypes.byref(out)) == A.ERR_INVALID_INPUT
    arr = (ctypes.c_void_p * 2)(None, None)
    assert lib.marl_ensemble_create(arr, 2, ctypes.byref(out)) == A.ERR_INVALID_INPUT
    assert lib.marl_ensemble_tick(None, None, None, 0, 1, None, None, None) == A.ERR_INVALID_INPUT
    assert lib.marl_ensemble_set_grid(None, 1) == A.ERR_INVALID_INPUT
    assert lib.marl_ensemble_launch_shape(None, None, None, None, None) == A.ERR_INVALID_INPUT
    assert lib.marl_ensemble_set_whole_columns(None, 1) == A.ERR_INVALID_INPUT
    assert lib.marl_set_graph_limits(None, 8, 8) == A.ERR_INVALID_INPUT
    assert lib.marl_set_cell_layout(None, A.CELL_LAYOUT_AUTO) == A.ERR_INVALID_INPUT
    assert lib.marl_time_field_phase(None, None, 1, 0, None) == A.ERR_INVALID_INPUT
    assert lib.marl_tick_graph(None, None, 0, 1, None, None) == A.ERR_INVALID_INPUT
    assert lib.marl_set_kinetics_mode(None, 0) == A.ERR_INVALID_INPUT
    assert lib.marl_surrogate_param_count() == 159516


def test_tile_table(pkg):
    """Tile ids: single-substep shapes first, then the two-substep (temporally
    blocked) ones; every id reports its shape."""
    lib = pkg.load_library()
    n = lib.marl_stencil_tile_count()
    subs = [lib.marl_stencil_tile_substeps(t) for t in range(n)]
    assert set(subs) == {1, 2} and subs == sorted(subs)
    assert lib.marl_stencil_tile_substeps(-1) < 0 and lib.marl_stencil_tile_substeps(n) < 0
    ry, zc, pd = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    sg, nw = ctypes.c_int32(), ctypes.c_int32()
    for t in range(n):
        assert lib.marl_stencil_tile_shape(t, ctypes.byref(ry), ctypes.byref(zc), ctypes.byref(pd)) == 0
        assert ry.value in (1, 2, 4) and zc.value in (4, 8, 16, 32, 64, 128) and pd.value in (1, 2)
        assert lib.marl_stencil_tile_split(t, ctypes.byref(sg), ctypes.byref(nw)) == 0
        assert nw.value in (4, 8, 12) and 12 % sg.value == 0
        if subs[t] == 1:  # a block's wave groups split the 12 species evenly
            assert 12 % (sg.value * nw.value // 4) == 0
    assert lib.marl_stencil_tile_split(n, None, None) != 0


def test_topology_hash_groups_wiring(pkg):
    """Host restatement of the surrogate's topology key: rulesets differing only
    in rates share a hash, any wiring change separates them."""
    from marl_ochre_amd import scenarios, starters

    r = np.zeros(4, pkg.RULESET_DTYPE)
    for k in range(4):
        r[k] = starters.make_ruleset(k % 3)
    r["reactions"]["v_max"][3] *= 2.0  # same wiring as ruleset 0
    h = scenarios.topology_hash(r)
    assert h[0] == h[3] and len(set(h[:3].tolist())) == 3
    r["reactions"]["substrate"][3, 5] = (int(r["reactions"]["substrate"][3, 5]) + 1) % 15
    assert scenarios.topology_hash(r)[3] != h[0]


def test_uploaded_records_carry_zero_padding(pkg):
    """np.concatenate packs the padding out of structured records (the default
    column is built that way).  The upload's conversion back to the C layout
    copies field by field into zeros, so padding and reserved bytes reach the
    device as zeros on every run; records already in the C layout pass through
    without a copy."""
    from marl_ochre_amd import engine, scenarios

    _, cells, rules, _ = scenarios.winogradsky_column((128, 128, 64), starter_state=True)
    A = pkg._abi
    assert cells.dtype != A.CELL_DTYPE and rules.dtype != A.RULESET_DTYPE   # packed by np.concatenate
    c = engine._records(cells, A.CELL_DTYPE)
    r = engine._records(rules, A.RULESET_DTYPE)
    assert c.dtype == A.CELL_DTYPE and r.dtype == A.RULESET_DTYPE
    cb = c.view(np.uint8).reshape(len(c), -1)
    rb = r.view(np.uint8).reshape(len(r), -1)
    named = np.zeros(A.RULESET_DTYPE.itemsize, bool)
    for name, (sub, off) in A.RULESET_DTYPE.fields.items():
        base, count = (sub.base, int(np.prod(sub.shape))) if sub.shape else (sub, 1)
        for k in range(count):
            o = off + k * base.itemsize
            if base.names:
                for fname, (fd, fo) in base.fields.items():
                    named[o + fo:o + fo + fd.itemsize] = True
            else:
                named[o:o + base.itemsize] = True
    assert (~named).sum() == 8 * 2 + 8 * 2 + 8   # transport and effector pads, reserved tail
    assert not cb[:, 86:].any() and not rb[:, ~named].any()
    for name in cells.dtype.names:
        assert np.array_equal(c[name], cells[name]), name
    for name in rules.dtype.names:
        sub = A.RULESET_DTYPE[name].base
        for f in (sub.names or [None]):
            got, want = (r[name], rules[name]) if f is None else (r[name][f], rules[name][f])
            assert np.array_equal(got, want), (name, f)
    assert engine._records(c, A.CELL_DTYPE) is c
