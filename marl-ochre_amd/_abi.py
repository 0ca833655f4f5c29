"""ctypes view of the C ABI declared in include/marl_hip.h.

The record layouts below are the byte layouts of marl_sim_params,
marl_cell_state and marl_ruleset; tests check their sizes against the header.
Loading fails loudly when libmarl_hip.so has not been built: there is no CPU
fallback in the product path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmarl_hip.so"

S_EXT, M_INT, R_MAX = 12, 16, 16
S_RECEPTORS = S_TRANSPORTERS = S_EFFECTORS = 8

OK, ERR_NO_DEVICE, ERR_DEVICE_INIT, ERR_TRANSFER, ERR_INVALID_INPUT = 0, 1, 2, 3, 4
EVENT_NONE, EVENT_DIVISION, EVENT_DEATH, EVENT_QUIESCENCE = 0, 1, 2, 3
KINETICS_PLAIN, KINETICS_SURROGATE, KINETICS_SHADOW = 0, 1, 2
CELL_LAYOUT_AUTO, CELL_LAYOUT_AOS, CELL_LAYOUT_QUADS = 0, 1, 2


class marl_dims(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32)]


class marl_sim_params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_float),
        ("diffusion_substeps", ctypes.c_int32),
        ("d_voxel", ctypes.c_float * S_EXT),
        ("lambda_decay", ctypes.c_float * S_EXT),
        ("source_rate_oxidant", ctypes.c_float),
        ("source_rate_carbon", ctypes.c_float),
        ("source_rate_reductant", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("c_max", ctypes.c_float),
        ("lambda_maintenance", ctypes.c_float),
        ("hard_death_floor", ctypes.c_float),
        ("reaction_maintenance", ctypes.c_float),
        ("base_division_prep", ctypes.c_float),
        ("prep_maintenance_multiplier", ctypes.c_float),
        ("rush_penalty_rate", ctypes.c_float),
        ("alpha_eps", ctypes.c_float),
        ("k_eps", ctypes.c_float),
        ("light_efficiency", ctypes.c_float),
        ("surface_intensity", ctypes.c_float),
        ("cell_absorption", ctypes.c_float),
        ("chemical_absorption", ctypes.c_float),
        ("light_floor", ctypes.c_float),
        ("mutation_stddev", ctypes.c_float),
        ("structural_mutation_rate_mult", ctypes.c_float),
        ("meta_mutation_rate", ctypes.c_float),
        ("meta_mutation_clamp_low", ctypes.c_float),
        ("meta_mutation_clamp_high", ctypes.c_float),
        ("hill_exponent_clamp_low", ctypes.c_float),
        ("hill_exponent_clamp_high", ctypes.c_float),
        ("active_reaction_threshold", ctypes.c_float),
        ("division_neighbor_distance", ctypes.c_int32),
        ("_pad0", ctypes.c_uint32),
        ("rng_seed", ctypes.c_uint64),
    ]


class marl_tick_stats(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int64),
        ("divisions", ctypes.c_int64),
        ("deaths", ctypes.c_int64),
        ("births_added", ctypes.c_int64),
        ("levels", ctypes.c_int32),
        ("_pad", ctypes.c_int32),
        ("ms_field", ctypes.c_float),
        ("ms_schedule", ctypes.c_float),
        ("ms_kinetics", ctypes.c_float),
        ("ms_fate", ctypes.c_float),
        ("ms_tick", ctypes.c_float),
        ("ms_transpose_back", ctypes.c_float),
    ]


class marl_train_config(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("batch", ctypes.c_int32),
        ("reinit", ctypes.c_int32),
        ("lr", ctypes.c_float),
        ("lr_final", ctypes.c_float),
        ("zmax", ctypes.c_float),
        ("energy_margin", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("max_pairs", ctypes.c_int64),
    ]


class marl_train_report(ctypes.Structure):
    _fields_ = [
        ("loss_first", ctypes.c_float),
        ("loss_last", ctypes.c_float),
        ("val_loss", ctypes.c_float),
        ("ms", ctypes.c_float),
        ("pairs", ctypes.c_int64),
        ("pairs_train", ctypes.c_int64),
    ]


class marl_surr_stats(ctypes.Structure):
    _fields_ = [
        ("cells", ctypes.c_int64),
        ("surrogate", ctypes.c_int64),
        ("fallback_topology", ctypes.c_int64),
        ("fallback_ood", ctypes.c_int64),
        ("fallback_margin", ctypes.c_int64),
        ("shadow", ctypes.c_int64),
        ("shadow_within_1e3", ctypes.c_int64),
        ("shadow_fate_agree", ctypes.c_int64),
        ("shadow_max_rel_err", ctypes.c_float),
        ("_pad", ctypes.c_int32),
        ("err_hist", ctypes.c_int64 * 24),
        ("shadow_within_1e3_species", ctypes.c_int64 * 16),
        ("fate_exact", ctypes.c_int64 * 4),
        ("fate_surrogate", ctypes.c_int64 * 4),
    ]


class marl_population_summary(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int64),
        ("sum_energy", ctypes.c_double),
        ("sum_enzyme_a", ctypes.c_double),
        ("sum_enzyme_b", ctypes.c_double),
        ("sum_active_reactions", ctypes.c_int64),
        ("z_thirds", ctypes.c_int64 * 3),
    ]


# surrogate network sizes (80 features -> 3 x 256 SiLU -> 28 targets)
SURR_FEAT, SURR_HIDDEN, SURR_TARG = 80, 256, 28

# numpy record layouts of marl_cell_state (96 B) and marl_ruleset (576 B)
CELL_DTYPE = np.dtype(
    {
        "names": ["internal", "pos", "quiescent", "starter_type", "lineage_id", "age", "prep_remaining"],
        "formats": [(np.float32, (M_INT,)), (np.uint16, (3,)), np.uint8, np.uint8, np.uint64, np.uint32, np.uint16],
        "offsets": [0, 64, 70, 71, 72, 80, 84],
        "itemsize": 96,
    }
)

_RECEPTOR = np.dtype({"names": ["k_half", "n_hill", "gain"], "formats": [np.float32] * 3, "offsets": [0, 4, 8],
                      "itemsize": 12})
_TRANSPORT = np.dtype({"names": ["uptake_rate", "secrete_rate", "ext_species", "int_species"],
                       "formats": [np.float32, np.float32, np.uint8, np.uint8], "offsets": [0, 4, 8, 9],
                       "itemsize": 12})
_REACTION = np.dtype({"names": ["substrate", "product", "catalyst", "cofactor", "k_m", "v_max", "k_cat"],
                      "formats": [np.uint8] * 4 + [np.float32] * 3, "offsets": [0, 1, 2, 3, 4, 8, 12],
                      "itemsize": 16})
_EFFECTOR = np.dtype({"names": ["threshold", "rate", "int_species", "ext_species"],
                      "formats": [np.float32, np.float32, np.uint8, np.uint8], "offsets": [0, 4, 8, 9],
                      "itemsize": 12})
RULESET_DTYPE = np.dtype(
    {
        "names": ["receptors", "transport", "reactions", "effectors", "division_energy", "death_energy",
                  "quiescence_energy", "division_prep_ticks", "hgt_propensity", "mutation_rate"],
        "formats": [(_RECEPTOR, (S_RECEPTORS,)), (_TRANSPORT, (S_TRANSPORTERS,)), (_REACTION, (R_MAX,)),
                    (_EFFECTOR, (S_EFFECTORS,))] + [np.float32] * 6,
        "offsets": [0, 96, 192, 448, 544, 548, 552, 556, 560, 564],
        "itemsize": 576,
    }
)


class GpuError(RuntimeError):
    """≙ GpuError (crates/marl-gpu/src/context.rs:4-9)."""

    code = -1


class NoAdapter(GpuError):
    code = ERR_NO_DEVICE


class DeviceRequest(GpuError):
    code = ERR_DEVICE_INIT


class BufferMap(GpuError):
    code = ERR_TRANSFER


class InvalidInput(GpuError):
    code = ERR_INVALID_INPUT


_ERRORS = {ERR_NO_DEVICE: NoAdapter, ERR_DEVICE_INIT: DeviceRequest, ERR_TRANSFER: BufferMap,
           ERR_INVALID_INPUT: InvalidInput}

_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

# name -> (restype, argtypes); every function declared in include/marl_hip.h
SIGNATURES = {
    "marl_create": (ctypes.c_int, [ctypes.POINTER(marl_dims), ctypes.POINTER(_P)]),
    "marl_destroy": (None, [_P]),
    "marl_last_error": (ctypes.c_char_p, [_P]),
    "marl_device_count": (ctypes.c_int, []),
    "marl_version": (ctypes.c_char_p, []),
    "marl_diffuse_tick": (ctypes.c_int, [_P, _P, _SZ, _P, _SZ, ctypes.POINTER(marl_sim_params)]),
    "marl_upload_field_aos": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_download_field_aos": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_upload_field_soa": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_download_field_soa": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_upload_occupancy": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_download_occupancy": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_apply_boundary_sources": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params)]),
    "marl_diffuse_substeps": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I32]),
    "marl_light_update": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params)]),
    "marl_download_light": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_upload_light": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_upload_cells": (ctypes.c_int, [_P, _P, _P, _SZ]),
    "marl_cell_count": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "marl_download_cells": (ctypes.c_int, [_P, _P, _P, _SZ, ctypes.POINTER(_I64)]),
    "marl_tick": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I64, _I32, ctypes.POINTER(marl_tick_stats)]),
    "marl_tick_graph": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I64, _I32,
                                       ctypes.POINTER(marl_tick_stats), ctypes.POINTER(_I32)]),
    "marl_set_kinetics_block": (ctypes.c_int, [_P, _I32]),
    "marl_get_kinetics_block": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "marl_autotune_kinetics": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), ctypes.POINTER(ctypes.c_int32)]),
    "marl_set_phase_budget": (ctypes.c_int, [_P, _I32]),
    "marl_tick_phase_budget": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_set_graph_limits": (ctypes.c_int, [_P, _I32, _I32]),
    "marl_set_cell_layout": (ctypes.c_int, [_P, _I32]),
    "marl_cell_pass": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _P, _SZ, ctypes.POINTER(_I32)]),
    "marl_mass_budget": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_time_diffuse": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I32, ctypes.POINTER(ctypes.c_float)]),
    "marl_time_light": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I32, ctypes.POINTER(ctypes.c_float)]),
    "marl_time_field_phase": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I32, _I32,
                                             ctypes.POINTER(ctypes.c_float)]),
    "marl_autotune_stencil": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), ctypes.POINTER(_I32)]),
    "marl_set_stencil_tile": (ctypes.c_int, [_P, _I32]),
    "marl_get_stencil_tile": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "marl_stencil_tile_count": (ctypes.c_int, []),
    "marl_stencil_tile_substeps": (ctypes.c_int, [_I32]),
    "marl_stencil_tile_shape": (ctypes.c_int, [_I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "marl_stencil_tile_split": (ctypes.c_int, [_I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "marl_copy_probe": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]),
    "marl_gather_probe": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]),
    "marl_mfma_probe": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]),
    "marl_hgt_transfer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]),
    "marl_seed_positions": (ctypes.c_int, [ctypes.POINTER(marl_dims), ctypes.c_uint64, ctypes.c_uint64, _I64, _I32,
                                           _I32, _I32, _P, _P, _P, _I64, ctypes.POINTER(_I64)]),
    "marl_sync": (ctypes.c_int, [_P]),
    "marl_cell_summary": (ctypes.c_int, [_P, ctypes.POINTER(marl_population_summary), _P, _SZ]),
    "marl_cell_summary_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(marl_population_summary), _P, _SZ]),
    "marl_chemistry_profile": (ctypes.c_int, [_P, _P, _SZ]),
    "marl_download_viewer_cells": (ctypes.c_int, [_P, _P, _SZ, ctypes.POINTER(_I64)]),
    "marl_sample": (ctypes.c_int, [_P, _P, _SZ, _P]),
    "marl_ensemble_create": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_P)]),
    "marl_ensemble_destroy": (None, [_P]),
    "marl_ensemble_last_error": (ctypes.c_char_p, [_P]),
    "marl_ensemble_tick": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _P, _I64, _I32, _P,
                                          ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_float)]),
    "marl_ensemble_set_tile": (ctypes.c_int, [_P, _I32]),
    "marl_ensemble_set_grid": (ctypes.c_int, [_P, _I32]),
    "marl_ensemble_time_field": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I32,
                                                ctypes.POINTER(ctypes.c_float)]),
    "marl_ensemble_autotune": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), ctypes.POINTER(_I32)]),
    "marl_ensemble_launch_shape": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                                                  ctypes.POINTER(_I32)]),
    "marl_ensemble_time_field_phase": (ctypes.c_int, [_P, ctypes.POINTER(marl_sim_params), _I32,
                                                      ctypes.POINTER(ctypes.c_float)]),
    "marl_ensemble_set_whole_columns": (ctypes.c_int, [_P, _I32]),
    "marl_ensemble_set_chunks": (ctypes.c_int, [_P, ctypes.POINTER(_I32), _I32]),
    "marl_ensemble_get_grid": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "marl_ensemble_get_chunks": (ctypes.c_int, [_P, ctypes.POINTER(_I32), _I32]),
    "marl_set_kinetics_mode": (ctypes.c_int, [_P, _I32]),
    "marl_harvest_begin": (ctypes.c_int, [_P, _I64, ctypes.c_float]),
    "marl_harvest_end": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "marl_harvest_download": (ctypes.c_int, [_P, _P, _P, _P, _SZ, ctypes.POINTER(_I64)]),
    "marl_harvest_upload": (ctypes.c_int, [_P, _P, _P, _P, _SZ]),
    "marl_surrogate_train": (ctypes.c_int, [_P, ctypes.POINTER(marl_train_config), ctypes.POINTER(marl_train_report)]),
    "marl_surrogate_param_count": (ctypes.c_int, []),
    "marl_surrogate_get_params": (ctypes.c_int, [_P, _P, _SZ, _P, _SZ]),
    "marl_surrogate_set_params": (ctypes.c_int, [_P, _P, _SZ, _P, _SZ]),
    "marl_surrogate_forward": (ctypes.c_int, [_P, _P, _SZ, _P, _I32]),
    "marl_surrogate_grad": (ctypes.c_int, [_P, _P, _P, _SZ, ctypes.POINTER(ctypes.c_float), _P]),
    "marl_surrogate_adam_step": (ctypes.c_int, [_P, _P, ctypes.c_float]),
    "marl_surrogate_configure": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "marl_surrogate_set_dead_zone": (ctypes.c_int, [_P, ctypes.c_float]),
    "marl_surrogate_stats": (ctypes.c_int, [_P, ctypes.POINTER(marl_surr_stats)]),
    "marl_surrogate_reset_stats": (ctypes.c_int, [_P]),
    "marl_surrogate_time_infer": (ctypes.c_int, [_P, _I64, _I32, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def load_library(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load libmarl_hip.so (built in-tree by the package Makefile)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `make -C marl-ochre_amd` (or __graft_entry__.build())")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def check(rc: int, ctx=None, lib=None, what: str = "") -> None:
    if rc == OK:
        return
    msg = what
    if ctx is not None and lib is not None:
        m = lib.marl_last_error(ctx)
        if m:
            msg = (what + ": " if what else "") + m.decode(errors="replace")
    raise _ERRORS.get(rc, GpuError)(msg or f"status {rc}")
