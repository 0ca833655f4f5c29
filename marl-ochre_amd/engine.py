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
            check(rc, what=msgs.get(rc, "failed to request GPU device"))
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

    def sync(self) -> None:
        self._chk(self.lib.marl_sync(self.h))

    # -- cells ----------------------------------------------------------------
    def upload_cells(self, cells: np.ndarray, rules: np.ndarray) -> None:
        c = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        r = np.ascontiguousarray(rules, dtype=RULESET_DTYPE)
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
