"""marl_hgt_transfer on the resident rulesets against the oracle's
hgt::transfer_reaction (crates/marl-cell/src/hgt.rs:7-27): bit-identical
rulesets, and the independence rules of a batch (state untouched on error)."""
import numpy as np
import pytest

from test_gpu_cells import population

pytestmark = pytest.mark.gpu


def test_hgt_matches_oracle(pkg, orc):
    A = pkg._abi
    dims = (64, 32, 24)
    cells, rules = population(A, dims, 6000, 12)
    rng = np.random.default_rng(3)
    full = rng.random(len(rules)) < 0.2                 # some recipients without an inactive slot
    rules["reactions"]["v_max"][full] = np.where(rules["reactions"]["v_max"][full] == 0, np.float32(0.3),
                                                 rules["reactions"]["v_max"][full])
    perm = rng.permutation(len(rules))
    pairs = np.stack([perm[:2000], perm[2000:4000]], 1)   # donors and recipients disjoint
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(np.zeros(np.prod(dims) * 12, np.float32))
        g.upload_cells(cells, rules)
        g.hgt_transfer(pairs, seed=99)
        c2, r2 = g.download_cells()
    want = orc.hgt_transfer(rules, pairs, 99)
    assert r2.tobytes() == want.tobytes()
    assert c2.tobytes() == cells.tobytes()
    assert (r2["reactions"].tobytes() != rules["reactions"].tobytes())


@pytest.mark.parametrize("pairs", [[(0, 1), (2, 1)], [(0, 1), (1, 2)], [(3, 3)], [(0, 10 ** 6)], [(-1, 2)]])
def test_hgt_rejects_dependent_or_invalid_pairs(pkg, pairs):
    A = pkg._abi
    dims = (32, 8, 8)
    cells, rules = population(A, dims, 50, 4)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(np.zeros(np.prod(dims) * 12, np.float32))
        g.upload_cells(cells, rules)
        with pytest.raises(pkg.InvalidInput):
            g.hgt_transfer(pairs, seed=1)
        assert g.download_cells()[1].tobytes() == rules.tobytes()
