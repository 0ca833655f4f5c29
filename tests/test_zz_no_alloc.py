"""marl_cell_summary_noalloc against marl_cell_summary through the C ABI.

The allocation-free variant keeps its partials and counters in a fixed device
array per GPU that keeps the last call's contents; two calls in a row, the
second on a smaller population and a shallower grid, must each equal the
allocating entry point bit for bit (FP64 sums, counts, per-layer counts)."""
import ctypes

import numpy as np
import pytest

from test_gpu_cells import population

pytestmark = pytest.mark.gpu


def _summary(g, fn):
    from marl_ochre_amd import _abi

    st = _abi.marl_population_summary()
    zc = np.full(g.dims[2], -7, np.int64)
    rc = getattr(g.lib, fn)(g.h, ctypes.byref(st), zc.ctypes.data, zc.size)
    assert rc == 0, (fn, rc, g.lib.marl_last_error(g.h))
    d = {k: getattr(st, k) for k, _ in st._fields_ if k != "z_thirds"}
    d["z_thirds"] = list(st.z_thirds)
    return d, zc


def _bits(d):
    return {k: (np.float64(v).tobytes() if isinstance(v, float) else v) for k, v in d.items()}


def test_noalloc_summary_matches_allocating_entry_point(pkg):
    A = pkg._abi
    cases = [((128, 96, 48), 20000, 1), ((64, 32, 16), 1500, 2), ((32, 8, 8), 0, 3)]
    for dims, n, seed in cases:
        print("summary case", dims, n, flush=True)
        cells, rules = population(A, dims, n, seed) if n else (np.zeros(0, A.CELL_DTYPE), np.zeros(0, A.RULESET_DTYPE))
        with pkg.MarlGpu(dims) as g:
            g.upload_cells(cells, rules)
            want, wz = _summary(g, "marl_cell_summary")
            got, gz = _summary(g, "marl_cell_summary_noalloc")
            assert _bits(got) == _bits(want), (dims, got, want)
            assert np.array_equal(gz, wz)
            assert got["n_cells"] == n and wz.sum() == n


def test_noalloc_summary_rejects_grids_above_its_limit(pkg):
    from marl_ochre_amd import _abi

    with pkg.MarlGpu((32, 2, 4100)) as g:
        st = _abi.marl_population_summary()
        rc = g.lib.marl_cell_summary_noalloc(g.h, ctypes.byref(st), None, 0)
        assert rc == _abi.ERR_INVALID_INPUT
        assert b"4096" in g.lib.marl_last_error(g.h)
