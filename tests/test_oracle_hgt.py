"""The oracle's hgt::transfer_reaction (crates/marl-cell/src/hgt.rs:7-27):
donor untouched, the recipient's first inactive slot (|v_max| < 1e-9) takes a
donor reaction with |v_max| > 1e-9, a drawn slot when none is inactive, no
change when the donor has no active reaction, and the drawn choices uniform
(the reference draws them from an OS-seeded rand ThreadRng, so parity with it
is at the level of these rules and distributions)."""
import numpy as np
import pytest
from scipy import stats


@pytest.fixture(scope="module")
def base(pkg):
    from marl_ochre_amd import starters

    return starters.make_ruleset(1)


def _rules(pkg, base, n):
    r = np.zeros(n, pkg.RULESET_DTYPE)
    r[:] = base
    return r


def _set_active(r, i, slots):
    r["reactions"]["v_max"][i] = 0.0
    for s in slots:
        r["reactions"]["v_max"][i, s] = np.float32(0.5 + s)
        r["reactions"]["substrate"][i, s] = s % 12
        r["reactions"]["product"][i, s] = (s + 3) % 12


def test_first_inactive_slot_and_donor_untouched(pkg, orc, base):
    r = _rules(pkg, base, 2)
    _set_active(r, 0, [2, 5, 9])
    _set_active(r, 1, [0, 1, 2, 3, 5])       # slot 4 is the first inactive one
    out = orc.hgt_transfer(r, [(0, 1)], seed=7)
    assert out[0].tobytes() == r[0].tobytes()
    changed = [q for q in range(16) if out["reactions"][1, q].tobytes() != r["reactions"][1, q].tobytes()]
    assert changed == [4]
    assert out["reactions"][1, 4].tobytes() in {r["reactions"][0, q].tobytes() for q in (2, 5, 9)}


def test_full_recipient_and_empty_donor(pkg, orc, base):
    r = _rules(pkg, base, 3)
    _set_active(r, 0, [7])
    _set_active(r, 1, range(16))              # no inactive slot: a drawn one is overwritten
    _set_active(r, 2, [])
    out = orc.hgt_transfer(r, [(0, 1)], seed=3)
    changed = [q for q in range(16) if out["reactions"][1, q].tobytes() != r["reactions"][1, q].tobytes()]
    assert len(changed) <= 1 and all(out["reactions"][1, q].tobytes() == r["reactions"][0, 7].tobytes()
                                     for q in changed)
    same = orc.hgt_transfer(r, [(2, 1)], seed=3)  # the donor has no active reaction
    assert same.tobytes() == r.tobytes()


def test_threshold_edges(pkg, orc, base):
    r = _rules(pkg, base, 3)
    _set_active(r, 0, [])
    r["reactions"]["v_max"][0, 3] = np.float32(1e-9)      # neither active (> 1e-9) ...
    _set_active(r, 1, [6])
    out = orc.hgt_transfer(r, [(0, 2)], seed=1)
    assert out.tobytes() == r.tobytes()                    # ... so the donor has nothing to give
    _set_active(r, 2, range(16))
    r["reactions"]["v_max"][2, 0] = np.float32(1e-9)       # ... nor inactive (< 1e-9): slot 0 is not the target
    r["reactions"]["v_max"][2, 11] = np.float32(-1e-12)    # |v_max| counts: slot 11 is the first inactive one
    out = orc.hgt_transfer(r, [(1, 2)], seed=1)
    changed = [q for q in range(16) if out["reactions"][2, q].tobytes() != r["reactions"][2, q].tobytes()]
    assert changed == [11] and out["reactions"][2, 11].tobytes() == r["reactions"][1, 6].tobytes()


def test_choices_are_uniform(pkg, orc, base):
    n = 6000
    r = _rules(pkg, base, 2 * n)
    donors = list(range(n))
    for i in donors:
        _set_active(r, i, [1, 4, 6, 10, 13])
    for i in range(n, 2 * n):
        _set_active(r, i, range(16))
        r["reactions"]["v_max"][i] += np.float32(100.0)   # every overwrite is visible
    out = orc.hgt_transfer(r, [(i, n + i) for i in range(n)], seed=11)
    picked, slots = [], []
    for i in range(n):
        diff = [q for q in range(16) if out["reactions"][n + i, q].tobytes() != r["reactions"][n + i, q].tobytes()]
        assert len(diff) == 1
        slots.append(diff[0])
        picked.append(int(np.round(out["reactions"][n + i, diff[0]]["v_max"] - 0.5)))
    _, p_pick = stats.chisquare(np.bincount(picked, minlength=16)[[1, 4, 6, 10, 13]])
    _, p_slot = stats.chisquare(np.bincount(slots, minlength=16))
    assert p_pick > 1e-4 and p_slot > 1e-4, (p_pick, p_slot)
