"""Writes the golden vectors under tests/golden/ from independent implementations.

philox4x64.json: numpy.random.Philox (the Random123 Philox4x64-10 bijection)
outputs for chosen counters and keys.  numpy increments the 256-bit counter
before producing a block, so a generator constructed with counter c - 1 yields
the block of counter c.
"""
import json
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).parent


def philox_cases():
    rng = np.random.default_rng(20261203)
    cases = []
    fixed = [([1, 0, 0, 0], [1, 2]), ([1, 0, 0, 0], [0, 0]), ([5, 77, 0, 0], [0, (2 << 32) | 3]),
             ([2**63, 1, 2, 3], [2**64 - 1, 12345])]
    for _ in range(12):
        key = [int(v) for v in rng.integers(0, 2**63, 2, dtype=np.uint64)]
        ctr = [int(v) for v in rng.integers(1, 2**63, 4, dtype=np.uint64)]
        fixed.append((ctr, key))
    for ctr, key in fixed:
        g = np.random.Philox(key=np.array(key, np.uint64), counter=np.array([ctr[0] - 1] + ctr[1:], np.uint64))
        cases.append({"ctr": ctr, "key": key, "out": [int(v) for v in g.random_raw(4)]})
    return cases


if __name__ == "__main__":
    (HERE / "philox4x64.json").write_text(json.dumps(philox_cases(), indent=1) + "\n")
    print("wrote", HERE / "philox4x64.json")
