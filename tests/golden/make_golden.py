"""Writes the golden vectors under tests/golden/ from independent implementations.

philox4x64.json: numpy.random.Philox (the Random123 Philox4x64-10 bijection)
outputs for chosen counters and keys.  numpy increments the 256-bit counter
before producing a block, so a generator constructed with counter c - 1 yields
the block of counter c.

mlp_torch.npz: the surrogate network (80 -> 256 -> 256 -> 256 -> 28, SiLU) in
torch float32 on the CPU with autograd: outputs, MSE loss, gradient and three
torch.optim.Adam steps, for parameters and inputs drawn by mlp_inputs(seed).
Gradient and Adam results are stored at every GRAD_STRIDE-th parameter plus
per-layer float64 norms, to keep the fixture small.
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


MLP_DIMS = (80, 256, 256, 256, 28)
MLP_SEED = 31
MLP_ROWS = 64
GRAD_STRIDE = 37
ADAM_LR = 1e-3
ADAM_SCALES = (1.0, 0.5, -1.0)  # gradient multipliers of the three Adam steps


def mlp_inputs(seed=MLP_SEED):
    """Flat float32 parameters (per layer W [in x out] row-major, then b), x and t."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in zip(MLP_DIMS[:-1], MLP_DIMS[1:]):
        lim = np.sqrt(6.0 / i)
        parts.append(rng.uniform(-lim, lim, (i, o)).astype(np.float32).reshape(-1))
        parts.append(rng.uniform(-0.1, 0.1, o).astype(np.float32))
    params = np.concatenate(parts)
    x = rng.standard_normal((MLP_ROWS, MLP_DIMS[0])).astype(np.float32)
    t = rng.standard_normal((MLP_ROWS, MLP_DIMS[-1])).astype(np.float32)
    return params, x, t


def mlp_case():
    import torch

    params, x, t = mlp_inputs()
    ws, off = [], 0
    for i, o in zip(MLP_DIMS[:-1], MLP_DIMS[1:]):
        W = torch.tensor(params[off:off + i * o].reshape(i, o), requires_grad=True)
        off += i * o
        b = torch.tensor(params[off:off + o], requires_grad=True)
        off += o
        ws += [W, b]
    h = torch.tensor(x)
    for l in range(4):
        h = h @ ws[2 * l] + ws[2 * l + 1]
        if l < 3:
            h = torch.nn.functional.silu(h)
    y = h
    loss = torch.mean((y - torch.tensor(t)) ** 2)
    loss.backward()
    grad = torch.cat([w.grad.reshape(-1) for w in ws]).detach().numpy().astype(np.float32)
    norms = np.array([np.linalg.norm(w.grad.numpy().astype(np.float64)) for w in ws])
    flat = torch.tensor(params.copy(), requires_grad=True)
    opt = torch.optim.Adam([flat], lr=ADAM_LR, betas=(0.9, 0.999), eps=1e-8)
    for sc in ADAM_SCALES:
        opt.zero_grad()
        flat.grad = torch.tensor(grad * np.float32(sc))
        opt.step()
    idx = np.arange(0, params.size, GRAD_STRIDE)
    return {"y": y.detach().numpy().astype(np.float32), "loss": np.float64(loss.item()), "grad_idx": idx,
            "grad": grad[idx], "grad_norms": norms, "adam": flat.detach().numpy()[idx].astype(np.float32),
            "torch_version": np.array(torch.__version__)}


if __name__ == "__main__":
    (HERE / "philox4x64.json").write_text(json.dumps(philox_cases(), indent=1) + "\n")
    print("wrote", HERE / "philox4x64.json")
    np.savez_compressed(HERE / "mlp_torch.npz", **mlp_case())
    print("wrote", HERE / "mlp_torch.npz")
