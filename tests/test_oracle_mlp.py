"""The surrogate network's NumPy oracle pinned against torch autograd goldens
(tests/golden/mlp_torch.npz) and float64 central differences (CPU only)."""
import importlib.util
import pathlib

import numpy as np
import pytest

from oracle import mlp_ref

GOLDEN = pathlib.Path(__file__).parent / "golden"


def _golden_module():
    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def case():
    mg = _golden_module()
    params, x, t = mg.mlp_inputs()
    g = dict(np.load(GOLDEN / "mlp_torch.npz"))
    return mg, params, x, t, g


def test_layout_size():
    assert mlp_ref.N_PARAMS == 80 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 28 + 28 == 159516


def test_forward_matches_torch(case):
    _, params, x, t, g = case
    y, _ = mlp_ref.forward(params, x)
    np.testing.assert_allclose(y, g["y"], rtol=1e-4, atol=2e-5)


def test_loss_and_grad_match_torch(case):
    _, params, x, t, g = case
    loss, grad = mlp_ref.loss_and_grad(params, x, t)
    assert abs(loss - float(g["loss"])) <= 1e-5 * abs(loss)
    np.testing.assert_allclose(grad[g["grad_idx"]], g["grad"], rtol=2e-3, atol=1e-8)
    lay, _ = mlp_ref.layout()
    norms = []
    for w, b, i, o in lay:
        norms += [np.linalg.norm(grad[w:w + i * o]), np.linalg.norm(grad[b:b + o])]
    np.testing.assert_allclose(norms, g["grad_norms"], rtol=1e-4)


def test_adam_matches_torch(case):
    mg, params, x, t, g = case
    _, grad = mlp_ref.loss_and_grad(params, x, t)
    # torch's gradient (float32) drives its Adam steps; use the oracle's, rounded to f32
    gt = grad.astype(np.float32)
    w, m, v = params.copy(), np.zeros_like(params), np.zeros_like(params)
    for step, sc in enumerate(mg.ADAM_SCALES, start=1):
        w, m, v = mlp_ref.adam(w, gt * np.float32(sc), m, v, step, mg.ADAM_LR)
    # Adam's first steps move each weight by about lr * sign(g): compare updates, not weights
    upd = w[g["grad_idx"]] - params[g["grad_idx"]]
    want = g["adam"] - params[g["grad_idx"]]
    np.testing.assert_allclose(upd, want, rtol=2e-3, atol=2e-7)


def test_grad_matches_central_differences(case):
    """float64 central differences of the float64 forward pin the hand-written
    chain rule independently of torch (a sign or transpose slip fails here)."""
    _, params, x, t, _ = case
    p64 = params.astype(np.float64)
    x, t = x[:8], t[:8]

    def loss64(p):
        layers, h = [], x.astype(np.float64)
        for w, b, i, o in mlp_ref.layout()[0]:
            layers.append((p[w:w + i * o].reshape(i, o), p[b:b + o]))
        for l, (W, b) in enumerate(layers):
            h = h @ W + b
            if l < 3:
                h = h / (1.0 + np.exp(-h))
        return np.mean((h - t) ** 2)

    # the oracle's gradient in float64 from the same parameters
    layers = mlp_ref.unpack(params)
    _, grad = mlp_ref.loss_and_grad(params, x, t)
    rng = np.random.default_rng(5)
    lay, total = mlp_ref.layout()
    picks = []
    for w, b, i, o in lay:  # a few weights and biases of every layer
        picks += list(rng.integers(w, w + i * o, 6)) + list(rng.integers(b, b + o, 2))
    h = 1e-5
    for k in picks:
        e = np.zeros(total)
        e[k] = h
        fd = (loss64(p64 + e) - loss64(p64 - e)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, grad[k])
    assert layers[0][0].shape == (80, 256)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5, 3.0e-39], np.float32)
    got = mlp_ref.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 ties to 1 + 2^-6 (even mantissa)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2**-6)
    assert got[3] == 1.0 and got[4] == -2.5
    assert (got.view(np.uint32) & 0xFFFF == 0).all()


def test_bf16_products_stay_close_to_exact(case):
    """Rounding every product operand to bf16 changes outputs by well under 5%."""
    _, params, x, _, _ = case
    y, _ = mlp_ref.forward(params, x)
    yb, _ = mlp_ref.forward(params, x, bf16_ops=True)
    rel = np.linalg.norm(yb - y) / np.linalg.norm(y)
    assert 1e-4 < rel < 5e-2
