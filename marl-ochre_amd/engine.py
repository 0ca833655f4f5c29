"""Python host mirror of the reference's GPU seam and of the resident tick.

`GpuFieldDiffuser` keeps the reference's names and contract
(crates/marl-gpu/src/field_diffusion.rs:31-268): `new()` fails with NoAdapter
when there is no GPU; `diffuse_tick_with_cells(field, occupancy, sim)` mutates
the AoS field in place, raises InvalidInput on wrong lengths, and is a no-op
when `sim.diffusion_substeps == 0`.

`MarlGpu` is the resident path: field, occupancy, light and cells stay in HBM
and only the calls' results cross PCIe.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import CELL_DTYPE, RULESET_DTYPE, S_EXT, check, ptr
from .config import GRID_X, GRID_Y, GRID_Z, SimulationConfig


def _params(sim) -> _abi.marl_sim_params:
    if isinstance(sim, _abi.marl_sim_params):
        return sim
    return sim.to_params()


def _records(a, dt: np.dtype) -> np.ndarray:
    """`a` as contiguous records of the C layout `dt`.  Records of another
    layout (numpy packs the padding out of structured arrays that pass through
    np.concatenate) are copied field by field into a zero-filled array, so the
    reserved and padding bytes the library stores are zeros, not whatever an
    uninitialised conversion buffer held."""
    a = np.asarray(a)
    if a.dtype == dt and a.flags.c_contiguous:
        return a
    out = np.zeros(a.shape, dt)
    for name in dt.names:
        out[name] = a[name]
    return out


class MarlGpu:
    """One GPU context (≙ GpuFieldDiffuser::new + the resident state)."""

    def __init__(self, dims=(GRID_X, GRID_Y, GRID_Z)):
        self.lib = _abi.load_library()
        self.dims = tuple(int(v) for v in dims)
        self.nvox = self.dims[0] * self.dims[1] * self.dims[2]
        d = _abi.marl_dims(*self.dims)
        h = ctypes.c_void_p()
        rc = self.lib.marl_create(ctypes.byref(d), ctypes.byref(h))
        if rc != _abi.OK:
            msgs = {
                _abi.ERR_NO_DEVICE: "no compatible GPU adapter found",
                _abi.ERR_INVALID_INPUT: f"unsupported grid {self.dims} (x must be a multiple of 32, "
                                        f"field under 2 GiB)",
            }
            why = self.lib.marl_last_error(None)
            what = msgs.get(rc, "failed to request GPU device")
            check(rc, what=what + (": " + why.decode(errors="replace") if why else ""))
        self.h = h

    # -- lifecycle --------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.marl_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str = "") -> None:
        check(rc, self.h, self.lib, what)

    # -- field ------------------------------------------------------------
    def upload_field_aos(self, field: np.ndarray) -> None:
        f = np.ascontiguousarray(field, dtype=np.float32).reshape(-1)
        self._chk(self.lib.marl_upload_field_aos(self.h, ptr(f), f.size))

    def download_field_aos(self) -> np.ndarray:
        out = np.empty(self.nvox * S_EXT, np.float32)
        self._chk(self.lib.marl_download_field_aos(self.h, ptr(out), out.size))
        return out

    def upload_field_soa(self, field: np.ndarray) -> None:
        f = np.ascontiguousarray(field, dtype=np.float32).reshape(-1)
        self._chk(self.lib.marl_upload_field_soa(self.h, ptr(f), f.size))

    def download_field_soa(self) -> np.ndarray:
        out = np.empty(self.nvox * S_EXT, np.float32)
        self._chk(self.lib.marl_download_field_soa(self.h, ptr(out), out.size))
        return out

    def upload_occupancy(self, occ: np.ndarray) -> None:
        o = np.ascontiguousarray(occ, dtype=np.uint8).reshape(-1)
        self._chk(self.lib.marl_upload_occupancy(self.h, ptr(o), o.size))

    def download_occupancy(self) -> np.ndarray:
        out = np.empty(self.nvox, np.uint8)
        self._chk(self.lib.marl_download_occupancy(self.h, ptr(out), out.size))
        return out

    def apply_boundary_sources(self, sim) -> None:
        p = _params(sim)
        self._chk(self.lib.marl_apply_boundary_sources(self.h, ctypes.byref(p)))

    def diffuse(self, sim, n_substeps: int | None = None) -> None:
        p = _params(sim)
        n = p.diffusion_substeps if n_substeps is None else n_substeps
        self._chk(self.lib.marl_diffuse_substeps(self.h, ctypes.byref(p), int(n)))

    def light_update(self, sim) -> None:
        p = _params(sim)
        self._chk(self.lib.marl_light_update(self.h, ctypes.byref(p)))

    def download_light(self) -> np.ndarray:
        out = np.empty(self.nvox, np.float32)
        self._chk(self.lib.marl_download_light(self.h, ptr(out), out.size))
        return out

    def upload_light(self, light: np.ndarray) -> None:
        lt = np.ascontiguousarray(light, dtype=np.float32).reshape(-1)
        self._chk(self.lib.marl_upload_light(self.h, ptr(lt), lt.size))

    # -- timing / tuning ----------------------------------------------------
    def time_diffuse(self, sim, n: int) -> float:
        p = _params(sim)
        ms = ctypes.c_float()
        self._chk(self.lib.marl_time_diffuse(self.h, ctypes.byref(p), int(n), ctypes.byref(ms)))
        return ms.value

    def time_field_phase(self, sim, n_ticks: int = 1, sources: bool = False) -> float:
        """n_ticks field phases (substeps, light fused where the tile allows)."""
        p = _params(sim)
        ms = ctypes.c_float()
        self._chk(self.lib.marl_time_field_phase(self.h, ctypes.byref(p), int(n_ticks), 1 if sources else 0,
                                                 ctypes.byref(ms)))
        return ms.value

    def time_light(self, sim, n: int) -> float:
        p = _params(sim)
        ms = ctypes.c_float()
        self._chk(self.lib.marl_time_light(self.h, ctypes.byref(p), int(n), ctypes.byref(ms)))
        return ms.value

    def autotune(self, sim) -> int:
        p = _params(sim)
        t = ctypes.c_int32()
        self._chk(self.lib.marl_autotune_stencil(self.h, ctypes.byref(p), ctypes.byref(t)))
        return t.value

    @property
    def tile(self) -> int:
        t = ctypes.c_int32()
        self._chk(self.lib.marl_get_stencil_tile(self.h, ctypes.byref(t)))
        return t.value

    @tile.setter
    def tile(self, t: int) -> None:
        self._chk(self.lib.marl_set_stencil_tile(self.h, int(t)))

    def tile_shape(self, t: int | None = None) -> tuple[int, int, int]:
        """(rows per thread, z-chunk, prefetch distance) of a stencil tile id."""
        t = self.tile if t is None else t
        ry, zc, pd = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.lib.marl_stencil_tile_shape(int(t), ctypes.byref(ry), ctypes.byref(zc), ctypes.byref(pd)))
        return ry.value, zc.value, pd.value

    def tile_split(self, t: int | None = None) -> tuple[int, int]:
        """(species per thread, waves per block) of a stencil tile id."""
        t = self.tile if t is None else t
        sg, nw = ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.lib.marl_stencil_tile_split(int(t), ctypes.byref(sg), ctypes.byref(nw)))
        return sg.value, nw.value

    def tile_substeps(self, t: int | None = None) -> int:
        """Substeps per launch of a tile (2 for the temporally blocked tiles)."""
        return int(self.lib.marl_stencil_tile_substeps(int(self.tile if t is None else t)))

    def sync(self) -> None:
        self._chk(self.lib.marl_sync(self.h))

    # -- cells ----------------------------------------------------------------
    def upload_cells(self, cells: np.ndarray, rules: np.ndarray) -> None:
        c = _records(cells, CELL_DTYPE)
        r = _records(rules, RULESET_DTYPE)
        if c.shape != r.shape:
            raise _abi.InvalidInput("cells and rulesets differ in length")
        self._chk(self.lib.marl_upload_cells(self.h, ptr(c), ptr(r), c.size))

    def cell_count(self) -> int:
        n = ctypes.c_int64()
        self._chk(self.lib.marl_cell_count(self.h, ctypes.byref(n)))
        return n.value

    def download_cells(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.cell_count()
        c = np.zeros(max(n, 1), CELL_DTYPE)
        r = np.zeros(max(n, 1), RULESET_DTYPE)
        got = ctypes.c_int64()
        self._chk(self.lib.marl_download_cells(self.h, ptr(c), ptr(r), c.size, ctypes.byref(got)))
        return c[: got.value], r[: got.value]

    def tick(self, sim, tick0: int = 0, n_ticks: int = 1) -> dict:
        p = _params(sim)
        st = _abi.marl_tick_stats()
        self._chk(self.lib.marl_tick(self.h, ctypes.byref(p), int(tick0), int(n_ticks), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_ if not k.startswith("_")}

    def tick_graph(self, sim, tick0: int = 0, n_ticks: int = 1) -> dict:
        """marl_tick through the captured hipGraph; 'host_finished' counts ticks
        whose cell phase exceeded the captured schedule budget."""
        p = _params(sim)
        st = _abi.marl_tick_stats()
        hf = ctypes.c_int32()
        self._chk(self.lib.marl_tick_graph(self.h, ctypes.byref(p), int(tick0), int(n_ticks), ctypes.byref(st),
                                           ctypes.byref(hf)))
        d = {k: getattr(st, k) for k, _ in st._fields_ if not k.startswith("_")}
        d["host_finished"] = hf.value
        return d

    CELL_LAYOUTS = {"auto": 0, "aos": 1, "quads": 2}

    def set_cell_layout(self, layout: str = "auto") -> None:
        """Field layout of the plain cell pass: "aos" (transposed view),
        "quads" (resident quads in place) or "auto"; results are identical."""
        self._chk(self.lib.marl_set_cell_layout(self.h, self.CELL_LAYOUTS[layout]))

    @property
    def kinetics_block(self) -> int:
        b = ctypes.c_int32()
        self._chk(self.lib.marl_get_kinetics_block(self.h, ctypes.byref(b)))
        return b.value

    @kinetics_block.setter
    def kinetics_block(self, block: int) -> None:
        self._chk(self.lib.marl_set_kinetics_block(self.h, int(block)))

    def autotune_kinetics(self, sim) -> int:
        """marl_autotune_kinetics: the fastest plain-pass block size on the
        resident state (left unchanged)."""
        p = _params(sim)
        b = ctypes.c_int32()
        self._chk(self.lib.marl_autotune_kinetics(self.h, ctypes.byref(p), ctypes.byref(b)))
        return b.value

    def set_graph_limits(self, sweeps: int = 0, levels: int = 0) -> None:
        """Schedule budget of the captured tick; (0, 0) = automatic (the default)."""
        self._chk(self.lib.marl_set_graph_limits(self.h, int(sweeps), int(levels)))

    def cell_pass(self, sim) -> tuple[np.ndarray, int]:
        p = _params(sim)
        n = self.cell_count()
        ev = np.zeros(max(n, 1), np.uint8)
        lv = ctypes.c_int32()
        self._chk(self.lib.marl_cell_pass(self.h, ctypes.byref(p), ptr(ev), ev.size, ctypes.byref(lv)))
        return ev[:n], lv.value

    def mass_budget(self) -> np.ndarray:
        out = np.zeros(40, np.float64)
        self._chk(self.lib.marl_mass_budget(self.h, ptr(out), out.size))
        return out

    # -- outputs ------------------------------------------------------------------
    def set_phase_budget(self, on: bool = True) -> None:
        self._chk(self.lib.marl_set_phase_budget(self.h, 1 if on else 0))

    def tick_phase_budget(self) -> np.ndarray:
        """[4, 40] FP64 budgets of the last tick: after the occupancy rebuild,
        the field phase, the cell pass, and fate with compaction."""
        out = np.zeros((4, 40), np.float64)
        self._chk(self.lib.marl_tick_phase_budget(self.h, ptr(out), out.size))
        return out

    def cell_summary(self) -> tuple[dict, np.ndarray]:
        """Population sums (FP64) and z-thirds, plus cells per z-layer."""
        st = _abi.marl_population_summary()
        zc = np.zeros(self.dims[2], np.int64)
        self._chk(self.lib.marl_cell_summary(self.h, ctypes.byref(st), ptr(zc), zc.size))
        d = {k: getattr(st, k) for k, _ in st._fields_ if k != "z_thirds"}
        d["z_thirds"] = list(st.z_thirds)
        return d, zc

    def chemistry_profile(self) -> np.ndarray:
        """[Z, 13]: per z-layer XY means of the 12 species, then of light (FP64)."""
        out = np.zeros((self.dims[2], S_EXT + 1), np.float64)
        self._chk(self.lib.marl_chemistry_profile(self.h, ptr(out), out.size))
        return out

    def download_viewer_cells(self) -> bytes:
        """The 25-byte packed viewer records of every cell, in cell order."""
        n = self.cell_count()
        buf = np.zeros(max(n, 1) * 25, np.uint8)
        got = ctypes.c_int64()
        self._chk(self.lib.marl_download_viewer_cells(self.h, ptr(buf), buf.size, ctypes.byref(got)))
        return buf[: got.value * 25].tobytes()

    def hgt_transfer(self, pairs, seed: int) -> None:
        """hgt::transfer_reaction for each (donor, recipient) row of pairs."""
        pr = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        self._chk(self.lib.marl_hgt_transfer(self.h, ptr(pr), len(pr), int(seed)))

    def sample(self, queries) -> np.ndarray:
        """Point reads: rows (x, y, z, species), species -1 = light."""
        q = np.ascontiguousarray(np.asarray(queries, np.int32).reshape(-1, 4))
        out = np.zeros(len(q), np.float32)
        self._chk(self.lib.marl_sample(self.h, ptr(q), len(q), ptr(out)))
        return out

    # -- learned kinetics surrogate --------------------------------------------
    def set_kinetics_mode(self, mode: int) -> None:
        self._chk(self.lib.marl_set_kinetics_mode(self.h, int(mode)))

    def harvest_begin(self, capacity: int, sample_rate: float = 1.0) -> None:
        self._chk(self.lib.marl_harvest_begin(self.h, int(capacity), float(sample_rate)))

    def harvest_end(self) -> int:
        n = ctypes.c_int64()
        self._chk(self.lib.marl_harvest_end(self.h, ctypes.byref(n)))
        return n.value

    def harvest_count(self) -> int:
        """Pairs kept so far (at most the capacity), without ending the harvest."""
        n = ctypes.c_int64()
        self._chk(self.lib.marl_harvest_download(self.h, None, None, None, 0, ctypes.byref(n)))
        return n.value

    def harvest_download(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = ctypes.c_int64()
        self._chk(self.lib.marl_harvest_download(self.h, None, None, None, 0, ctypes.byref(n)))
        m = max(n.value, 1)
        f = np.zeros((m, _abi.SURR_FEAT), np.float32)
        t = np.zeros((m, _abi.SURR_TARG), np.float32)
        h = np.zeros(m, np.uint64)
        self._chk(self.lib.marl_harvest_download(self.h, ptr(f), ptr(t), ptr(h), m, ctypes.byref(n)))
        return f[: n.value], t[: n.value], h[: n.value]

    def harvest_upload(self, feat: np.ndarray, targ: np.ndarray, hashes: np.ndarray) -> None:
        f = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(targ, np.float32)
        h = np.ascontiguousarray(hashes, np.uint64)
        if f.shape != (h.size, _abi.SURR_FEAT) or t.shape != (h.size, _abi.SURR_TARG):
            raise _abi.InvalidInput("pair arrays have inconsistent shapes")
        self._chk(self.lib.marl_harvest_upload(self.h, ptr(f), ptr(t), ptr(h), h.size))

    def surrogate_train(self, steps: int, batch: int = 4096, lr: float = 2e-3, lr_final: float = 1e-5,
                        seed: int = 1, reinit: bool = False, zmax: float = 6.0, energy_margin: float = 0.05,
                        max_pairs: int = 0) -> dict:
        cfg = _abi.marl_train_config(steps=int(steps), batch=int(batch), reinit=int(reinit), lr=lr,
                                     lr_final=lr_final, zmax=zmax, energy_margin=energy_margin, seed=int(seed),
                                     max_pairs=int(max_pairs))
        rep = _abi.marl_train_report()
        self._chk(self.lib.marl_surrogate_train(self.h, ctypes.byref(cfg), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_ if not k.startswith("_")}

    @property
    def surrogate_param_count(self) -> int:
        return int(self.lib.marl_surrogate_param_count())

    def surrogate_get_params(self) -> tuple[np.ndarray, np.ndarray]:
        w = np.zeros(self.surrogate_param_count, np.float32)
        nm = np.zeros(2 * (_abi.SURR_FEAT + _abi.SURR_TARG), np.float32)
        self._chk(self.lib.marl_surrogate_get_params(self.h, ptr(w), w.size, ptr(nm), nm.size))
        return w, nm

    def surrogate_set_params(self, params: np.ndarray, norm: np.ndarray | None = None) -> None:
        w = np.ascontiguousarray(params, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._chk(self.lib.marl_surrogate_set_params(self.h, ptr(w), w.size, None if nm is None else ptr(nm),
                                                     0 if nm is None else nm.size))

    def surrogate_forward(self, x: np.ndarray, fused: bool = True) -> np.ndarray:
        xx = np.ascontiguousarray(x, np.float32)
        if xx.ndim != 2 or xx.shape[1] != _abi.SURR_FEAT:
            raise _abi.InvalidInput("x must be [rows, 80]")
        y = np.zeros((xx.shape[0], _abi.SURR_TARG), np.float32)
        self._chk(self.lib.marl_surrogate_forward(self.h, ptr(xx), xx.shape[0], ptr(y), int(fused)))
        return y

    def surrogate_grad(self, x: np.ndarray, t: np.ndarray) -> tuple[float, np.ndarray]:
        xx = np.ascontiguousarray(x, np.float32)
        tt = np.ascontiguousarray(t, np.float32)
        if xx.ndim != 2 or xx.shape[1] != _abi.SURR_FEAT or tt.shape != (xx.shape[0], _abi.SURR_TARG):
            raise _abi.InvalidInput("x must be [rows, 80] and t [rows, 28]")
        g = np.zeros(self.surrogate_param_count, np.float32)
        loss = ctypes.c_float()
        self._chk(self.lib.marl_surrogate_grad(self.h, ptr(xx), ptr(tt), xx.shape[0], ctypes.byref(loss), ptr(g)))
        return loss.value, g

    def surrogate_adam_step(self, grads: np.ndarray, lr: float) -> None:
        g = np.ascontiguousarray(grads, np.float32)
        if g.size != self.surrogate_param_count:
            raise _abi.InvalidInput("gradient size mismatch")
        self._chk(self.lib.marl_surrogate_adam_step(self.h, ptr(g), float(lr)))

    def surrogate_configure(self, zmax: float = 6.0, energy_margin: float = 0.05, shadow_rate: float = 0.0) -> None:
        self._chk(self.lib.marl_surrogate_configure(self.h, float(zmax), float(energy_margin), float(shadow_rate)))

    def surrogate_set_dead_zone(self, dead_zone: float) -> None:
        """Drop corrections of at most `dead_zone` target-scale units (the
        explicit step stands there)."""
        self._chk(self.lib.marl_surrogate_set_dead_zone(self.h, float(dead_zone)))

    def surrogate_stats(self) -> dict:
        st = _abi.marl_surr_stats()
        self._chk(self.lib.marl_surrogate_stats(self.h, ctypes.byref(st)))
        arrays = ("err_hist", "shadow_within_1e3_species", "fate_exact", "fate_surrogate")
        d = {k: getattr(st, k) for k, _ in st._fields_ if not k.startswith("_") and k not in arrays}
        for k in arrays:
            d[k] = list(getattr(st, k))
        return d

    def surrogate_reset_stats(self) -> None:
        self._chk(self.lib.marl_surrogate_reset_stats(self.h))

    def surrogate_time_infer(self, rows: int, reps: int = 10) -> float:
        """Milliseconds per fused inference launch over `rows` synthetic rows."""
        ms = ctypes.c_float()
        self._chk(self.lib.marl_surrogate_time_infer(self.h, int(rows), int(reps), ctypes.byref(ms)))
        return ms.value


class GpuFieldDiffuser:
    """≙ marl_gpu::GpuFieldDiffuser (crates/marl-gpu/src/field_diffusion.rs:20-268)."""

    def __init__(self, dims=(GRID_X, GRID_Y, GRID_Z)):
        self._gpu = MarlGpu(dims)

    @classmethod
    def new(cls, dims=(GRID_X, GRID_Y, GRID_Z)) -> "GpuFieldDiffuser":
        return cls(dims)

    def diffuse_tick_with_cells(self, field: np.ndarray, occupancy: np.ndarray, sim: SimulationConfig) -> None:
        """Run sim.diffusion_substeps masked substeps on the AoS field, in place."""
        if field.dtype != np.float32 or not field.flags.c_contiguous:
            raise _abi.InvalidInput("field must be a C-contiguous float32 array")
        occ = np.ascontiguousarray(occupancy, dtype=np.uint8).reshape(-1)
        p = _params(sim)
        g = self._gpu
        g._chk(g.lib.marl_diffuse_tick(g.h, ptr(field), field.size, ptr(occ), occ.size, ctypes.byref(p)))

    def close(self) -> None:
        self._gpu.close()


class MarlEnsemble:
    """Independent columns of uneven sizes and seeds advanced together: one
    hipGraph per tick with every column's field phase batched into single
    launches and the cell phases side by side.  The columns are MarlGpu
    contexts the caller keeps (upload, download and close them as usual;
    close the ensemble first)."""

    def __init__(self, columns):
        self.lib = _abi.load_library()
        self.columns = list(columns)
        arr = (ctypes.c_void_p * len(self.columns))(*[c.h.value for c in self.columns])
        h = ctypes.c_void_p()
        rc = self.lib.marl_ensemble_create(arr, len(self.columns), ctypes.byref(h))
        check(rc, what="marl_ensemble_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.marl_ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int) -> None:
        if rc != _abi.OK:
            m = self.lib.marl_ensemble_last_error(self.h)
            raise _abi._ERRORS.get(rc, _abi.GpuError)(m.decode(errors="replace") if m else f"status {rc}")

    def tick(self, sim, seeds=None, tick0: int = 0, n_ticks: int = 1):
        """Returns (per-column stats dicts, GPU ms of the graphs, host-finished count)."""
        p = _params(sim)
        n = len(self.columns)
        st = (_abi.marl_tick_stats * n)()
        sd = None if seeds is None else np.ascontiguousarray(seeds, np.uint64)
        hf = ctypes.c_int32()
        ms = ctypes.c_float()
        self._chk(self.lib.marl_ensemble_tick(self.h, ctypes.byref(p), None if sd is None else ptr(sd), int(tick0),
                                              int(n_ticks), st, ctypes.byref(hf), ctypes.byref(ms)))
        stats = [{k: getattr(s, k) for k, _ in s._fields_ if not k.startswith("_")} for s in st]
        return stats, ms.value, hf.value

    def set_tile(self, tile: int) -> None:
        self._chk(self.lib.marl_ensemble_set_tile(self.h, int(tile)))

    def set_grid(self, mode) -> None:
        """Batched stencil grid: persistent (False / 0), one block per tile
        (True / 1) or persistent over balanced whole-depth pieces (2)."""
        self._chk(self.lib.marl_ensemble_set_grid(self.h, int(mode)))

    def grid_mode(self) -> int:
        m = ctypes.c_int32()
        self._chk(self.lib.marl_ensemble_get_grid(self.h, ctypes.byref(m)))
        return m.value

    def time_field(self, sim, n_substeps: int = 20) -> float:
        p = _params(sim)
        ms = ctypes.c_float()
        self._chk(self.lib.marl_ensemble_time_field(self.h, ctypes.byref(p), int(n_substeps), ctypes.byref(ms)))
        return ms.value

    def autotune(self, sim) -> int:
        p = _params(sim)
        t = ctypes.c_int32()
        self._chk(self.lib.marl_ensemble_autotune(self.h, ctypes.byref(p), ctypes.byref(t)))
        return t.value

    def launch_shape(self) -> dict:
        """The batched stencil launch: tile id, grid blocks, logical tiles, z-chunk mode."""
        t, g, n, w = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.lib.marl_ensemble_launch_shape(self.h, ctypes.byref(t), ctypes.byref(g), ctypes.byref(n),
                                                      ctypes.byref(w)))
        return {"tile": t.value, "grid_blocks": g.value, "stencil_tiles": n.value, "whole_columns": bool(w.value),
                "grid_mode": self.grid_mode()}

    def time_field_phase(self, sim, n_ticks: int = 5) -> float:
        """Milliseconds per batched field phase (substeps + light) of a tick."""
        p = _params(sim)
        ms = ctypes.c_float()
        self._chk(self.lib.marl_ensemble_time_field_phase(self.h, ctypes.byref(p), int(n_ticks), ctypes.byref(ms)))
        return ms.value

    def set_whole_columns(self, whole: bool) -> None:
        """Every column one z-chunk deep, the light scan fused into the last substep."""
        self._chk(self.lib.marl_ensemble_set_whole_columns(self.h, 1 if whole else 0))

    def set_chunks(self, chunks=None) -> None:
        """Per-column z-chunk planes (0: the tile shape's); None clears."""
        if chunks is None:
            self._chk(self.lib.marl_ensemble_set_chunks(self.h, None, 0))
            return
        a = (ctypes.c_int32 * len(chunks))(*[int(z) for z in chunks])
        self._chk(self.lib.marl_ensemble_set_chunks(self.h, a, len(chunks)))

    def chunks(self) -> list:
        """The z-chunk each column's batched stencil tiles use."""
        n = len(self.columns)
        a = (ctypes.c_int32 * n)()
        self._chk(self.lib.marl_ensemble_get_chunks(self.h, a, n))
        return list(a)
