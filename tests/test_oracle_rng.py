"""Pins the oracle's Philox4x64-10, its draw mapping, expf/logf and normals (CPU only)."""
import json
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).parent / "golden" / "philox4x64.json"


def test_philox_matches_numpy_golden_vectors(orc):
    cases = json.loads(GOLDEN.read_text())
    assert len(cases) >= 8
    for c in cases:
        got = orc.philox(c["ctr"], c["key"])
        assert [int(v) for v in got] == c["out"], c


def test_philox_matches_numpy_live(orc):
    rng = np.random.default_rng(11)
    for _ in range(20):
        key = [int(v) for v in rng.integers(0, 2**63, 2)]
        ctr = [int(v) for v in rng.integers(1, 2**63, 4)]
        g = np.random.Philox(key=np.array(key, np.uint64), counter=np.array([ctr[0] - 1] + ctr[1:], np.uint64))  # numpy bumps the counter first
        assert [int(v) for v in g.random_raw(4)] == [int(v) for v in orc.philox(ctr, key)]


def test_draw_layout(orc):
    # draw j = word (j>>1)&3 of block j>>3 of counter (j>>3, entity, 0, 0), low half for even j
    seed, stream, ent = 123, (2 << 32) | 9, 77
    for j in range(0, 40, 3):
        blk = orc.philox([j >> 3, ent, 0, 0], [seed, stream])
        w = int(blk[(j >> 1) & 3])
        want = (w >> 32) if (j & 1) else (w & 0xFFFFFFFF)
        assert orc.draw_u32(seed, stream, ent, j) == want


def _ulp(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


def test_expf_within_2ulp_of_libm(orc):
    xs = np.concatenate([np.linspace(-87.0, 88.0, 4001), np.linspace(-1, 1, 2001), [0.0, -0.5, -0.3, 1e-8]])
    worst = max(_ulp(orc.expf(x), np.exp(np.float32(x))) for x in xs.astype(np.float32))
    assert worst <= 2
    assert orc.expf(-100.0) == 0.0 and np.isinf(orc.expf(100.0))


def test_logf_within_2ulp_of_libm(orc):
    xs = np.geomspace(1e-30, 1e30, 4001).astype(np.float32)
    worst = max(_ulp(orc.logf(x), np.log(x)) for x in xs if abs(np.log(x)) > 1e-3)
    assert worst <= 2
    near1 = np.linspace(0.9, 1.1, 2001).astype(np.float32)
    assert max(abs(float(orc.logf(x)) - np.log(np.float64(x))) for x in near1) < 2e-7


def test_normal_sampler_moments(orc):
    import ctypes

    vals = []
    draws = ctypes.c_uint64(0)
    for e in range(4000):
        d = np.zeros(1, np.uint64)
        vals.append(orc.L.orc_normal(5, 7, e, d.ctypes.data, 1.0))
    v = np.array(vals, np.float64)
    assert abs(v.mean()) < 0.06 and abs(v.std() - 1.0) < 0.05
    assert abs(((v - v.mean()) ** 3).mean()) < 0.15
    del draws
