"""ctypes binding of the CPU oracle (liboracle.so) — test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to
check and time against; never by the product path.  Record layouts come from
the product's _abi module (types only).
"""
from __future__ import annotations

import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
ROOT = ORACLE_DIR.parent
LIB = ORACLE_DIR / "liboracle.so"


def _abi():
    sys.path.insert(0, str(ROOT))
    try:
        from __graft_entry__ import load_package
    finally:
        sys.path.pop(0)
    return load_package()._abi


def build() -> None:
    subprocess.run(["make", "-B", "-C", str(ORACLE_DIR)], check=True, capture_output=True)


def build_planted(plant: int, out: Path) -> Path:
    """The oracle built with planted mistake `plant` (ORC_PLANT in marl_oracle.c),
    for tests that show a check notices that mistake."""
    subprocess.run(["make", "-C", str(ORACLE_DIR), "plant", f"PLANT={int(plant)}", f"PLANT_OUT={out}"], check=True,
                   capture_output=True)
    return out


class Oracle:
    def __init__(self, lib: Path | None = None):
        if lib is None and not LIB.exists():
            build()
        A = _abi()
        self.A = A
        L = ctypes.CDLL(str(lib or LIB))
        P, D, SP = ctypes.c_void_p, ctypes.POINTER(A.marl_dims), ctypes.POINTER(A.marl_sim_params)
        u64, i64, f32 = ctypes.c_uint64, ctypes.c_int64, ctypes.c_float
        sig = {
            "orc_philox4x64": (None, [P, P, P]),
            "orc_draw_u32": (ctypes.c_uint32, [u64, u64, u64, u64]),
            "orc_expf": (f32, [f32]),
            "orc_logf": (f32, [f32]),
            "orc_apply_boundary_sources": (None, [P, D, SP]),
            "orc_diffusion_substep": (None, [P, P, P, D, SP, f32, ctypes.c_int]),
            "orc_diffuse_tick": (None, [P, P, P, D, SP, ctypes.c_int]),
            "orc_diffusion_substep_f64": (None, [P, P, P, D, SP, ctypes.c_double]),
            "orc_field_ticks_f64": (None, [P, P, P, D, SP, ctypes.c_int]),
            "orc_light_update": (None, [P, P, P, D, SP]),
            "orc_light_update_f64": (None, [P, P, P, D, SP]),
            "orc_cell_tick": (ctypes.c_int, [P, P, P, f32, SP, P]),
            "orc_mutate": (None, [P, SP, u64, u64, u64, P]),
            "orc_normal": (f32, [u64, u64, u64, P, f32]),
            "orc_hgt_transfer": (None, [P, P, i64, u64]),
            "orc_find_empty_neighbor": (ctypes.c_int, [P, P, D, SP, u64, u64, u64, P, P]),
            "orc_compact_indices": (i64, [i64, P, i64, P]),
            "orc_tick": (ctypes.c_int, [D, SP, u64, P, P, P, P, P, P, i64, P, P, ctypes.c_int]),
            "orc_cell_pass": (None, [D, SP, P, P, P, P, i64, P]),
            "orc_seed_positions": (i64, [D, u64, u64, i64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, P, P,
                                         i64]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        self.L = L

    # helpers ---------------------------------------------------------------
    @staticmethod
    def _p(a):
        return ctypes.c_void_p(a.ctypes.data)

    def _dims(self, dims):
        return ctypes.byref(self.A.marl_dims(*dims))

    @staticmethod
    def _sim(sim):
        p = sim if not hasattr(sim, "to_params") else sim.to_params()
        return ctypes.byref(p), p

    # RNG / math ---------------------------------------------------------------
    def philox(self, ctr, key):
        c = np.array(ctr, np.uint64)
        k = np.array(key, np.uint64)
        o = np.zeros(4, np.uint64)
        self.L.orc_philox4x64(self._p(c), self._p(k), self._p(o))
        return o

    def draw_u32(self, seed, stream, entity, draw):
        return self.L.orc_draw_u32(seed, stream, entity, draw)

    def expf(self, x):
        return np.float32(self.L.orc_expf(float(x)))

    def logf(self, x):
        return np.float32(self.L.orc_logf(float(x)))

    # field ------------------------------------------------------------------
    def boundary(self, field, dims, sim):
        f = np.ascontiguousarray(field, np.float32).copy()
        s, _keep = self._sim(sim)
        self.L.orc_apply_boundary_sources(self._p(f), self._dims(dims), s)
        return f

    def diffuse_tick(self, field, occ, dims, sim, nthreads=8):
        f = np.ascontiguousarray(field, np.float32).copy()
        scratch = np.empty_like(f)
        o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_diffuse_tick(self._p(f), self._p(scratch), None if o is None else self._p(o), self._dims(dims), s,
                                nthreads)
        return f

    def substep(self, field, occ, dims, sim, dt_sub, nthreads=8):
        f = np.ascontiguousarray(field, np.float32)
        out = np.empty_like(f)
        o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_diffusion_substep(self._p(f), self._p(out), None if o is None else self._p(o), self._dims(dims), s,
                                     float(dt_sub), nthreads)
        return out

    def substep_f64(self, field, occ, dims, sim, dt_sub):
        f = np.ascontiguousarray(field, np.float64)
        out = np.empty_like(f)
        o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_diffusion_substep_f64(self._p(f), self._p(out), None if o is None else self._p(o),
                                         self._dims(dims), s, float(dt_sub))
        return out

    def field_ticks_f64(self, field64, occ, dims, sim, n_ticks):
        """Advance the float64 field in place by n_ticks field-only ticks."""
        assert field64.dtype == np.float64 and field64.flags.c_contiguous
        scratch = np.empty_like(field64)
        o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_field_ticks_f64(self._p(field64), self._p(scratch), None if o is None else self._p(o),
                                   self._dims(dims), s, int(n_ticks))
        return field64

    def light(self, field, occ, dims, sim):
        f = np.ascontiguousarray(field, np.float32)
        out = np.empty(dims[0] * dims[1] * dims[2], np.float32)
        o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_light_update(self._p(out), self._p(f), None if o is None else self._p(o), self._dims(dims), s)
        return out

    def light_f64(self, field, occ, dims, sim):
        f = np.ascontiguousarray(field, np.float64)
        out = np.empty(dims[0] * dims[1] * dims[2], np.float64)
        o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_light_update_f64(self._p(out), self._p(f), None if o is None else self._p(o), self._dims(dims), s)
        return out

    # cells --------------------------------------------------------------------
    def cell_tick(self, cell, rule, ext, light, sim):
        c = np.array(cell, dtype=self.A.CELL_DTYPE).reshape(1)
        r = np.array(rule, dtype=self.A.RULESET_DTYPE).reshape(1)
        e = np.ascontiguousarray(ext, np.float32)
        d = np.zeros(12, np.float32)
        s, _keep = self._sim(sim)
        ev = self.L.orc_cell_tick(self._p(c), self._p(r), self._p(e), float(light), s, self._p(d))
        return c[0], d, ev

    def mutate(self, rule, sim, seed, stream, entity, draws=0):
        r = np.array(rule, dtype=self.A.RULESET_DTYPE).reshape(1)
        dr = np.array([draws], np.uint64)
        s, _keep = self._sim(sim)
        self.L.orc_mutate(self._p(r), s, seed, stream, entity, self._p(dr))
        return r[0], int(dr[0])

    def hgt_transfer(self, rules, pairs, seed):
        """Rulesets after hgt::transfer_reaction for each (donor, recipient) pair."""
        src = np.ascontiguousarray(rules, dtype=self.A.RULESET_DTYPE)
        r = np.empty_like(src)
        r.view(np.uint8)[:] = src.view(np.uint8)  # byte copy: padding included
        pr = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        self.L.orc_hgt_transfer(self._p(r), self._p(pr), len(pr), seed)
        return r

    def find_empty_neighbor(self, pos, cell_at, dims, sim, seed, stream, entity, draws=0):
        """(target (x, y, z) or None, draws after) of find_empty_neighbor."""
        p = np.ascontiguousarray(pos, np.uint16).reshape(3)
        ca = np.ascontiguousarray(cell_at, np.int32)
        d = np.array([draws], np.uint64)
        out = np.zeros(3, np.uint16)
        s, _keep = self._sim(sim)
        ok = self.L.orc_find_empty_neighbor(self._p(p), self._p(ca), self._dims(dims), s, seed, stream, entity,
                                            self._p(d), self._p(out))
        return (tuple(int(v) for v in out) if ok else None), int(d[0])

    def compact_indices(self, n, deaths):
        d = np.ascontiguousarray(sorted(deaths), np.int64)
        perm = np.zeros(max(n, 1), np.int64)
        m = self.L.orc_compact_indices(n, self._p(d) if d.size else None, d.size, self._p(perm))
        return perm[:m]

    def cell_pass(self, dims, sim, field, light, cells, rules):
        f = np.ascontiguousarray(field, np.float32).copy()
        lt = np.ascontiguousarray(light, np.float32)
        c = np.array(cells, dtype=self.A.CELL_DTYPE).copy()
        r = np.ascontiguousarray(rules, dtype=self.A.RULESET_DTYPE)
        ev = np.zeros(max(c.size, 1), np.uint8)
        s, _keep = self._sim(sim)
        self.L.orc_cell_pass(self._dims(dims), s, self._p(f), self._p(lt), self._p(c), self._p(r), c.size,
                             self._p(ev))
        return f, c, ev[: c.size]

    def tick(self, dims, sim, tick, field, cells, rules, nthreads=8):
        """One full tick; returns (field, light, cells, rules, events, stats)."""
        f = np.ascontiguousarray(field, np.float32).copy()
        scratch = np.empty_like(f)
        nv = dims[0] * dims[1] * dims[2]
        lt = np.zeros(nv, np.float32)
        n0 = len(cells)
        cap = 2 * n0 + 16
        c = np.zeros(cap, self.A.CELL_DTYPE)
        r = np.zeros(cap, self.A.RULESET_DTYPE)
        c[:n0] = cells
        r[:n0] = rules
        n = ctypes.c_int64(n0)
        ev = np.zeros(max(n0, 1), np.uint8)
        st = np.zeros(4, np.int64)
        s, _keep = self._sim(sim)
        rc = self.L.orc_tick(self._dims(dims), s, tick, self._p(f), self._p(scratch), self._p(lt), self._p(c),
                             self._p(r), ctypes.byref(n), cap, self._p(ev), self._p(st), nthreads)
        if rc != 0:
            raise RuntimeError("oracle tick capacity exceeded")
        stats = dict(zip(["divisions", "deaths", "births_added", "n_cells"], (int(v) for v in st)))
        return f, lt, c[: n.value].copy(), r[: n.value].copy(), ev[:n0], stats

    def seed_positions(self, dims, seed, band, count, margin, z_lo, z_hi, cell_at=None, index_base=0):
        nv = dims[0] * dims[1] * dims[2]
        ca = np.full(nv, -1, np.int32) if cell_at is None else cell_at
        pos = np.zeros((max(count, 1), 3), np.uint16)
        lin = np.zeros(max(count, 1), np.uint64)
        m = self.L.orc_seed_positions(self._dims(dims), seed, band, count, margin, z_lo, z_hi, self._p(ca),
                                      self._p(pos), self._p(lin), index_base)
        return pos[:m], lin[:m], ca
