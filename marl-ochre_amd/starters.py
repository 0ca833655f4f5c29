"""The three starter metabolisms as (cell, ruleset) records.

Values restated from crates/marl-sim/src/starter_metabolisms.rs:6-459
(inactive defaults :6-42; phototroph :46-187; chemolithotroph :191-326;
anaerobe :330-459).
"""
from __future__ import annotations

import numpy as np

from ._abi import CELL_DTYPE, RULESET_DTYPE

PHOTOTROPH, CHEMOLITHOTROPH, ANAEROBE = 0, 1, 2


def _inactive() -> np.ndarray:
    r = np.zeros((), RULESET_DTYPE)
    r["receptors"]["k_half"] = 1.0
    r["receptors"]["n_hill"] = 2.0
    r["receptors"]["gain"] = 0.0
    r["transport"]["uptake_rate"] = 0.0
    r["transport"]["secrete_rate"] = 0.0
    r["transport"]["ext_species"] = 0
    r["transport"]["int_species"] = 0
    rx = r["reactions"]
    rx["substrate"] = rx["product"] = rx["catalyst"] = 0
    rx["cofactor"] = 0xFF
    rx["k_m"] = 1.0
    rx["v_max"] = 0.0
    rx["k_cat"] = 0.5
    r["effectors"]["threshold"] = 10.0
    r["effectors"]["rate"] = 0.0
    r["effectors"]["int_species"] = 0
    r["effectors"]["ext_species"] = 0
    r["hgt_propensity"] = 0.1
    r["mutation_rate"] = 0.05
    return r


def _set(arr, i, **kw):
    for k, v in kw.items():
        arr[i][k] = v


def _rxn(r, i, sub, prod, cat, cof, km, vmax, kcat):
    _set(r["reactions"], i, substrate=sub, product=prod, catalyst=cat, cofactor=cof, k_m=km, v_max=vmax, k_cat=kcat)


def make_ruleset(kind: int) -> np.ndarray:
    r = _inactive()
    rc, tp, ef = r["receptors"], r["transport"], r["effectors"]
    if kind == PHOTOTROPH:
        _set(rc, 2, k_half=0.3, n_hill=2.0, gain=1.0)
        _set(rc, 3, k_half=0.5, n_hill=2.0, gain=1.0)
        _set(tp, 0, uptake_rate=0.6, secrete_rate=0.0, ext_species=3, int_species=3)
        _set(tp, 1, uptake_rate=0.0, secrete_rate=0.6, ext_species=1, int_species=1)
        _set(tp, 2, uptake_rate=0.0, secrete_rate=0.3, ext_species=4, int_species=4)
        _set(tp, 3, uptake_rate=0.1, secrete_rate=0.0, ext_species=2, int_species=2)
        _rxn(r, 0, 3, 0, 15, 0xFF, 0.2, 0.8, 0.1)
        _rxn(r, 1, 3, 1, 15, 0xFF, 0.2, 0.4, 0.1)
        _rxn(r, 2, 3, 4, 0, 0xFF, 0.3, 0.3, 0.3)
        _rxn(r, 3, 3, 5, 6, 0xFF, 0.3, 0.15, 0.1)
        _rxn(r, 4, 3, 6, 5, 0xFF, 0.3, 0.1, 0.1)
        _set(ef, 0, threshold=0.5, rate=0.8, int_species=1, ext_species=1)
        _set(ef, 1, threshold=1.0, rate=0.2, int_species=4, ext_species=4)
        fate = (1.2, 0.05, 0.15, 20.0)
    elif kind == CHEMOLITHOTROPH:
        _set(rc, 1, k_half=0.3, n_hill=2.0, gain=1.0)
        _set(rc, 2, k_half=0.3, n_hill=2.0, gain=1.0)
        _set(tp, 0, uptake_rate=0.7, secrete_rate=0.0, ext_species=1, int_species=1)
        _set(tp, 1, uptake_rate=0.7, secrete_rate=0.0, ext_species=2, int_species=2)
        _set(tp, 2, uptake_rate=0.2, secrete_rate=0.0, ext_species=3, int_species=3)
        _set(tp, 3, uptake_rate=0.0, secrete_rate=0.3, ext_species=4, int_species=4)
        _rxn(r, 0, 2, 0, 5, 1, 0.15, 0.7, 0.1)
        _rxn(r, 1, 1, 4, 0, 0xFF, 0.3, 0.3, 0.2)
        _rxn(r, 2, 7, 0, 5, 0xFF, 0.3, 0.15, 0.1)
        _rxn(r, 3, 3, 5, 6, 0xFF, 0.3, 0.15, 0.1)
        _rxn(r, 4, 3, 6, 5, 0xFF, 0.3, 0.1, 0.1)
        _set(ef, 0, threshold=0.5, rate=0.3, int_species=4, ext_species=4)
        fate = (1.0, 0.05, 0.12, 20.0)
    elif kind == ANAEROBE:
        _set(rc, 2, k_half=0.3, n_hill=2.0, gain=1.0)
        _set(tp, 0, uptake_rate=0.9, secrete_rate=0.0, ext_species=2, int_species=2)
        _set(tp, 1, uptake_rate=0.3, secrete_rate=0.0, ext_species=3, int_species=3)
        _set(tp, 2, uptake_rate=0.0, secrete_rate=0.5, ext_species=4, int_species=4)
        _set(tp, 3, uptake_rate=0.1, secrete_rate=0.0, ext_species=1, int_species=1)
        _rxn(r, 0, 2, 0, 5, 0xFF, 0.15, 0.8, 0.1)
        _rxn(r, 1, 3, 4, 0, 0xFF, 0.2, 0.3, 0.2)
        _rxn(r, 2, 0, 3, 1, 0xFF, 0.01, 2.0, 0.01)
        _rxn(r, 3, 3, 5, 6, 0xFF, 0.3, 0.15, 0.1)
        _rxn(r, 4, 3, 6, 5, 0xFF, 0.3, 0.1, 0.1)
        _set(ef, 0, threshold=0.5, rate=0.5, int_species=4, ext_species=4)
        fate = (0.8, 0.05, 0.1, 20.0)
    else:
        raise ValueError(kind)
    (r["division_energy"], r["death_energy"], r["quiescence_energy"], r["division_prep_ticks"]) = fate
    return r


_START_INTERNAL = {
    PHOTOTROPH: {0: 0.3, 3: 0.3, 5: 0.05, 6: 0.05},
    CHEMOLITHOTROPH: {0: 1.5, 1: 0.5, 2: 0.5, 5: 0.05, 6: 0.05, 7: 5.0},
    ANAEROBE: {0: 0.5, 2: 0.7, 5: 0.05, 6: 0.05},
}


def make_cell(kind: int, pos, lineage_id: int) -> np.ndarray:
    c = np.zeros((), CELL_DTYPE)
    c["pos"] = pos
    c["lineage_id"] = lineage_id
    for k, v in _START_INTERNAL[kind].items():
        c["internal"][k] = v
    c["starter_type"] = kind
    return c


def make_cells(kinds: np.ndarray, pos: np.ndarray, lineage_ids: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Vectorised make_cell / make_ruleset: (cells, rulesets) of the given starter kinds."""
    kinds = np.asarray(kinds)
    n = kinds.size
    cells = np.zeros(n, CELL_DTYPE)
    rules = np.zeros(n, RULESET_DTYPE)
    cells["pos"] = pos
    cells["lineage_id"] = lineage_ids
    cells["starter_type"] = kinds
    for kind in (PHOTOTROPH, CHEMOLITHOTROPH, ANAEROBE):
        sel = kinds == kind
        if not sel.any():
            continue
        rules[sel] = make_ruleset(kind)
        internal = np.zeros(16, np.float32)
        for k, v in _START_INTERNAL[kind].items():
            internal[k] = v
        cells["internal"][sel] = internal
    return cells, rules
