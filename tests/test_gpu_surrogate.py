"""GPU checks of the learned kinetics surrogate through the C ABI.

* network kernels: the BF16 MFMA forward (training GEMM path and fused
  inference), the backward and Adam, against oracle/mlp_ref.py (which is
  pinned to torch autograd goldens and float64 central differences);
* harvest: pairs the plain kinetics pass appends agree with the oracle's cell
  pass for the same cells;
* fallback and shadow: cells the surrogate does not take are advanced by the
  plain path bit for bit; shadow mode leaves the state bit-identical to plain;
* training: the network trained on harvested pairs predicts held-out cells
  better than the untrained network and than "no reaction" (zero deltas).
"""
import importlib.util
import pathlib

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu

F = np.float32
GOLDEN = pathlib.Path(__file__).parent / "golden"


def _golden():
    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def net():
    mg = _golden()
    params, x, t = mg.mlp_inputs()
    return mg, params, x, t


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("rows", [64, 1, 200, 700])
def test_forward_matches_oracle(pkg, net, rows):
    _, params, x0, _ = net
    rng = np.random.default_rng(rows)
    x = x0 if rows == 64 else rng.standard_normal((rows, 80)).astype(F)
    emu, _ = mlp_ref.forward(params, x, bf16_ops=True)
    exact, _ = mlp_ref.forward(params, x)
    with pkg.MarlGpu((32, 32, 32)) as g:
        g.surrogate_set_params(params)
        yf = g.surrogate_forward(x, fused=True)
        yt = g.surrogate_forward(x, fused=False)
    scale = max(1.0, float(np.abs(emu).max()))
    # against the bf16-operand restatement: only f32 summation order and exp differ
    assert np.abs(yf - emu).max() <= 4e-3 * scale, np.abs(yf - emu).max()
    assert np.abs(yt - emu).max() <= 4e-3 * scale, np.abs(yt - emu).max()
    assert np.abs(yf - yt).max() <= 2e-3 * scale
    # and the cost of bf16 operands against the exact network
    assert _rel(yf, exact) < 3e-2


@pytest.mark.parametrize("rows", [64, 200])
def test_backward_matches_oracle(pkg, net, rows):
    _, params, x0, t0 = net
    rng = np.random.default_rng(100 + rows)
    x = x0 if rows == 64 else rng.standard_normal((rows, 80)).astype(F)
    t = t0 if rows == 64 else rng.standard_normal((rows, 28)).astype(F)
    l_emu, g_emu = mlp_ref.loss_and_grad(params, x, t, bf16_ops=True)
    l_ex, g_ex = mlp_ref.loss_and_grad(params, x, t)
    with pkg.MarlGpu((32, 32, 32)) as g:
        g.surrogate_set_params(params)
        loss, grad = g.surrogate_grad(x, t)
    assert abs(loss - l_emu) <= 2e-3 * l_emu
    for w, b, i, o in mlp_ref.layout()[0]:
        for lo, hi in ((w, w + i * o), (b, b + o)):
            assert _rel(grad[lo:hi], g_emu[lo:hi]) < 1e-2, (lo, _rel(grad[lo:hi], g_emu[lo:hi]))
            assert _rel(grad[lo:hi], g_ex[lo:hi]) < 6e-2, (lo, _rel(grad[lo:hi], g_ex[lo:hi]))
    cos = float(grad @ g_ex / (np.linalg.norm(grad) * np.linalg.norm(g_ex)))
    assert cos > 0.998


def test_backward_matches_torch_golden(pkg, net):
    _, params, x, t = net
    gold = dict(np.load(GOLDEN / "mlp_torch.npz"))
    with pkg.MarlGpu((32, 32, 32)) as g:
        g.surrogate_set_params(params)
        loss, grad = g.surrogate_grad(x, t)
    assert abs(loss - float(gold["loss"])) <= 1e-2 * float(gold["loss"])
    idx = gold["grad_idx"]
    assert _rel(grad[idx], gold["grad"]) < 6e-2


def test_adam_matches_oracle(pkg, net):
    mg, params, x, t = net
    _, grad = mlp_ref.loss_and_grad(params, x, t)
    grad = grad.astype(F)
    w, m, v = params.copy(), np.zeros_like(params), np.zeros_like(params)
    with pkg.MarlGpu((32, 32, 32)) as g:
        g.surrogate_set_params(params)
        for step, sc in enumerate(mg.ADAM_SCALES, start=1):
            g.surrogate_adam_step(grad * F(sc), mg.ADAM_LR)
            w, m, v = mlp_ref.adam(w, grad * F(sc), m, v, step, mg.ADAM_LR)
        got, _ = g.surrogate_get_params()
    np.testing.assert_allclose(got - params, w - params, rtol=1e-4, atol=1e-9)


def _population(A, dims, n, seed, shared=True):
    from marl_ochre_amd import scenarios
    from marl_ochre_amd import starters as st

    rng = np.random.default_rng(seed)
    nv = dims[0] * dims[1] * dims[2]
    vox = np.sort(rng.choice(nv, n, replace=False))
    kinds = rng.integers(0, 3, n)
    cells = np.zeros(n, A.CELL_DTYPE)
    rules = np.zeros(n, A.RULESET_DTYPE)
    for k in range(3):
        rules[kinds == k] = st.make_ruleset(k)
    cells["pos"] = np.stack([vox % dims[0], (vox // dims[0]) % dims[1], vox // (dims[0] * dims[1])], 1)
    cells["starter_type"] = kinds
    cells["lineage_id"] = rng.integers(0, 2**62, n)
    cells["internal"] = rng.random((n, 16), dtype=F)
    scenarios.random_reactions(rules, rng, (1e-3, 1.0), (1e-3, 1.0), light_p=0.2, groups=kinds if shared else None)
    field = rng.random(nv * 12, dtype=F)
    light = rng.random(nv, dtype=F)
    return cells, rules, field, light


def test_harvest_pairs_match_oracle(pkg, orc):
    """Harvested pairs: the state features as each cell saw them, the explicit
    step restated on the host from those inputs (bit for bit), and targets equal
    to the oracle's post-kinetics internals minus the pre-tick ones minus the
    explicit deltas, divided by each output's target scale (its gross explicit
    activity plus 1e-5; 0, with target 0, for an output no explicit term
    touches) and clipped to +-4."""
    from marl_ochre_amd import scenarios

    A = pkg._abi
    dims = (64, 32, 32)
    cells, rules, field, light = _population(A, dims, 3000, 7, shared=False)
    sim = pkg.SimulationConfig()
    _, c_w, _ = orc.cell_pass(dims, sim, field, light, cells, rules)
    with pkg.MarlGpu(dims) as g:
        g.upload_field_aos(field)
        g.upload_cells(cells, rules)
        g.upload_light(light)
        g.harvest_begin(10000, 1.0)
        g.cell_pass(sim)
        n = g.harvest_end()
        feat, targ, hsh = g.harvest_download()
    assert n == len(cells)
    # pairs arrive in completion order: match them to cells by their internal state
    key = {cells["internal"][i].tobytes()[:60]: i for i in range(len(cells))}
    idx = np.array([key[feat[j, :15].astype(F).tobytes()] for j in range(n)])
    assert np.unique(idx).size == n
    np.testing.assert_array_equal(hsh, scenarios.topology_hash(rules)[idx])
    np.testing.assert_array_equal(feat[:, 15], cells["quiescent"][idx].astype(F))
    np.testing.assert_array_equal(feat[:, 29], (cells["prep_remaining"][idx] > 0).astype(F))
    vox = (cells["pos"][:, 2].astype(np.int64) * dims[1] + cells["pos"][:, 1]) * dims[0] + cells["pos"][:, 0]
    np.testing.assert_array_equal(feat[:, 28], light[vox[idx]])
    np.testing.assert_array_equal(feat[:, 73:], 0)
    for j in range(0, n, 7):   # every 7th pair through the host restatement
        i = idx[j]
        dI, dD, frel, gI, gD = mlp_ref.explicit_step(cells["internal"][i], feat[j, 16:28], feat[j, 28], rules[i],
                                                     bool(cells["quiescent"][i]), int(cells["prep_remaining"][i]),
                                                     sim, gross=True)
        np.testing.assert_array_equal(feat[j, 30:45], dI[:15])
        np.testing.assert_array_equal(feat[j, 45:57], dD)
        np.testing.assert_array_equal(feat[j, 57:73], frel)
        want = mlp_ref.scaled_targets((c_w["internal"][i] - cells["internal"][i]) - dI,
                                      mlp_ref.target_scales(gI, gD)[:16])
        np.testing.assert_array_equal(targ[j, :15], want[:15])
        assert targ[j, 15] == 0


@pytest.fixture(scope="module")
def trained(pkg):
    """A context whose network was trained on pairs harvested from two plain
    passes over 200k cells, each with its own random reaction wiring, plus the
    same network before training (parameters and normalisation)."""
    A = pkg._abi
    dims = (128, 128, 64)
    cells, rules, field, light = _population(A, dims, 200000, 11, shared=False)
    sim = pkg.SimulationConfig()
    g = pkg.MarlGpu(dims)
    g.upload_field_aos(field)
    g.upload_light(light)
    g.upload_cells(cells, rules)
    g.harvest_begin(500000, 1.0)
    for _ in range(2):
        g.cell_pass(sim)
    n_pairs = g.harvest_end()
    g.surrogate_train(steps=0, batch=4096, seed=3, reinit=True)
    untrained = g.surrogate_get_params()
    rep = g.surrogate_train(steps=3000, batch=4096, lr=2e-3, seed=3)
    yield dict(g=g, rep=rep, n_pairs=n_pairs, untrained=untrained, dims=dims, sim=sim)
    g.close()


def _held_out(pkg, dims, n, seed):
    """A fresh population with fresh per-cell wiring."""
    return _population(pkg._abi, dims, n, seed, shared=False)


def test_training_beats_untrained_and_explicit_step(pkg, trained):
    """Held-out cells of unseen wiring, in target-scale units (residual over
    each output's gross explicit activity): the internal-delta RMS error of
    explicit step + trained residual against the explicit step alone
    (residual 0) and against explicit step + untrained network."""
    g, rep = trained["g"], trained["rep"]
    assert trained["n_pairs"] == 400000 and rep["pairs"] == 400000
    assert rep["loss_last"] < 0.5 * rep["loss_first"]
    assert rep["val_loss"] < 0.5, rep   # standardised targets: predicting the mean scores 1.0
    c2, r2, f2, l2 = _held_out(pkg, trained["dims"], 20000, 12)
    g.upload_field_aos(f2)
    g.upload_light(l2)
    g.upload_cells(c2, r2)
    g.harvest_begin(30000, 1.0)
    g.cell_pass(trained["sim"])
    g.harvest_end()
    feat, targ, _ = g.harvest_download()

    def rms_err(w, norm):
        with pkg.MarlGpu((32, 32, 32)) as h:
            h.surrogate_set_params(w, norm)
            z = (feat - norm[:80]) / norm[80:160]
            pred = h.surrogate_forward(z, fused=True) * norm[188:216] + norm[160:188]
        return np.sqrt(np.mean((pred[:, :15] - targ[:, :15]) ** 2))

    err_net = rms_err(*g.surrogate_get_params())
    err_untrained = rms_err(*trained["untrained"])
    err_explicit = np.sqrt(np.mean(targ[:, :15] ** 2))   # the targets are the explicit step's error
    print("rms internal-delta error: net", err_net, "explicit step", err_explicit, "untrained", err_untrained)
    assert err_net < 0.97 * err_explicit, (err_net, err_explicit)
    assert err_net < 0.8 * err_untrained, (err_net, err_untrained)


def test_operating_point_keeps_explicit_accuracy_and_fates(pkg, trained):
    """The shipped operating point (dead zone 0.2) in shadow mode on held-out
    cells, against the explicit step alone (a dead zone no correction
    passes): per species, the share of cells within 1e-3 stays within 15 % of
    the explicit step's, the fate decisions agree on >= 95 % of cells, and no
    more cells land above a relative error of 1 than under the explicit step
    alone, within 25 %."""
    A = pkg._abi
    g, sim = trained["g"], trained["sim"]
    c2, r2, f2, l2 = _held_out(pkg, trained["dims"], 50000, 14)

    def shadow(dead_zone):
        g.upload_field_aos(f2)
        g.upload_light(l2)
        g.upload_cells(c2, r2)
        g.surrogate_configure(zmax=6.0, energy_margin=0.05, shadow_rate=0.0)
        g.surrogate_set_dead_zone(dead_zone)
        g.surrogate_reset_stats()
        g.set_kinetics_mode(A.KINETICS_SHADOW)
        try:
            g.cell_pass(sim)
            return g.surrogate_stats()
        finally:
            g.set_kinetics_mode(A.KINETICS_PLAIN)
            g.surrogate_set_dead_zone(0.2)

    op, ex = shadow(0.2), shadow(1e30)
    assert op["shadow"] == ex["shadow"] > 0.5 * len(c2)
    n = op["shadow"]
    w_op = np.array(op["shadow_within_1e3_species"][:15]) / n
    w_ex = np.array(ex["shadow_within_1e3_species"][:15]) / n
    print("within 1e-3 per species: operating point", np.round(w_op, 3), "explicit step", np.round(w_ex, 3))
    assert (w_op >= 0.85 * w_ex).all(), (w_op, w_ex)
    assert op["shadow_fate_agree"] >= 0.95 * n
    big_op, big_ex = sum(op["err_hist"][18:]), sum(ex["err_hist"][18:])   # bins from relative error 1 up
    print("cells above relative error 1: operating point", big_op, "explicit step", big_ex)
    assert big_op <= 1.25 * big_ex + 0.01 * n, (big_op, big_ex)


def test_surrogate_mode_accounting_and_bounds(pkg, trained):
    A = pkg._abi
    g, sim = trained["g"], trained["sim"]
    c2, r2, f2, l2 = _held_out(pkg, trained["dims"], 50000, 13)
    g.upload_field_aos(f2)
    g.upload_light(l2)
    g.upload_cells(c2, r2)
    g.surrogate_configure(zmax=6.0, energy_margin=0.05, shadow_rate=0.5)
    g.surrogate_reset_stats()
    g.set_kinetics_mode(A.KINETICS_SURROGATE)
    try:
        for _ in range(3):
            g.cell_pass(sim)
        st = g.surrogate_stats()
        c, _ = g.download_cells()
        f = g.download_field_aos()
    finally:
        g.set_kinetics_mode(A.KINETICS_PLAIN)
    n = 3 * len(c2)
    assert st["cells"] == n
    assert st["surrogate"] + st["fallback_topology"] + st["fallback_ood"] + st["fallback_margin"] == n
    assert st["fallback_topology"] == 0
    assert st["surrogate"] > 0.5 * n
    assert 0.45 * st["surrogate"] < st["shadow"] < 0.55 * st["surrogate"]
    assert sum(st["err_hist"]) == st["shadow"] and st["shadow_within_1e3"] <= st["shadow"]
    assert 0.9 * st["shadow"] <= st["shadow_fate_agree"] <= st["shadow"]
    assert all(st["shadow_within_1e3"] <= w <= st["shadow"] for w in st["shadow_within_1e3_species"][:15])
    assert sum(st["fate_exact"]) == sum(st["fate_surrogate"]) == st["shadow"]
    ci = c["internal"]
    assert np.isfinite(ci).all() and (ci >= 0).all() and (ci <= sim.c_max).all()
    assert np.isfinite(f).all() and (f >= 0).all()


def test_fallback_and_shadow_are_bit_exact(pkg):
    from marl_ochre_amd import scenarios

    A = pkg._abi
    dims = (64, 48, 32)
    cells, rules, field, light = _population(A, dims, 12000, 21)
    sim = pkg.SimulationConfig()

    def run(mode, zmax=1e9, margin=0.05, shadow=0.0):
        with pkg.MarlGpu(dims) as g:
            g.upload_field_aos(field)
            g.upload_light(light)
            g.upload_cells(cells, rules)
            if mode != A.KINETICS_PLAIN:
                rng = np.random.default_rng(0)
                g.surrogate_set_params(rng.standard_normal(g.surrogate_param_count).astype(F) * F(0.05), None)
                g.surrogate_configure(zmax=zmax, energy_margin=margin, shadow_rate=shadow)
                g.set_kinetics_mode(mode)
            ev, _ = g.cell_pass(sim)
            c, r = g.download_cells()
            st = g.surrogate_stats()
            return ev, c, g.download_field_aos(), st

    ev0, c0, f0, _ = run(A.KINETICS_PLAIN)
    # a vanishing out-of-distribution bound: every cell falls back
    ev1, c1, f1, st1 = run(A.KINETICS_SURROGATE, zmax=1e-20)
    assert st1["cells"] == len(cells) and st1["fallback_ood"] == len(cells) and st1["surrogate"] == 0
    assert st1["fallback_topology"] == 0
    # an energy margin wider than any energy: every cell falls back
    ev2, c2, f2, st2 = run(A.KINETICS_SURROGATE, margin=1e9)
    assert st2["fallback_margin"] == len(cells) and st2["surrogate"] == 0
    # shadow: exact path authoritative, every eligible cell compared
    ev3, c3, f3, st3 = run(A.KINETICS_SHADOW)
    assert st3["shadow"] == len(cells) and st3["surrogate"] == 0
    assert sum(st3["err_hist"]) == st3["shadow"]
    for ev, c, f in ((ev1, c1, f1), (ev2, c2, f2), (ev3, c3, f3)):
        np.testing.assert_array_equal(ev, ev0)
        np.testing.assert_array_equal(c["internal"].view(np.uint32), c0["internal"].view(np.uint32))
        np.testing.assert_array_equal(c["prep_remaining"], c0["prep_remaining"])
        np.testing.assert_array_equal(f.view(np.uint32), f0.view(np.uint32))
