"""NumPy restatement of the kinetics surrogate's network — test infrastructure only.

The surrogate has no counterpart in the reference (it replaces the reaction
phase of CellState::tick, crates/marl-cell/src/cell.rs:223-329, for eligible
cells).  This module is the checker for the HIP network kernels:

* ``forward`` — 80 -> 256 -> 256 -> 256 -> 28, SiLU on the hidden layers, in
  float64; with ``bf16=True`` every matrix-product operand is rounded to
  bfloat16 (round to nearest even) first, as the MFMA kernels do, while sums,
  biases and activations stay wide;
* ``loss_and_grad`` — mean squared error over all outputs and its gradient by
  the chain rule (hand-written, pinned against torch autograd goldens and
  float64 central differences in tests/test_oracle_mlp.py), with the same
  optional operand rounding;
* ``adam`` — Kingma & Ba (2015), float32 elementwise, bias-corrected moments.

The flat parameter vector holds, per layer, W [in x out] row-major, then b.
"""
from __future__ import annotations

import numpy as np

DIMS = (80, 256, 256, 256, 28)


def layout():
    """[(w_offset, b_offset, n_in, n_out)] per layer and the total length."""
    out, off = [], 0
    for l in range(4):
        i, o = DIMS[l], DIMS[l + 1]
        out.append((off, off + i * o, i, o))
        off += i * o + o
    return out, off


N_PARAMS = layout()[1]


def unpack(params: np.ndarray):
    lay, total = layout()
    assert params.size == total
    return [(params[w:w + i * o].reshape(i, o), params[b:b + o]) for w, b, i, o in lay]


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers]).astype(np.float32)


def bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bfloat16, returned as float32 (finite inputs)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(a))


def _op(a, rounded):
    return bf16(a).astype(np.float64) if rounded else np.asarray(a, np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params: np.ndarray, x: np.ndarray, bf16_ops: bool = False):
    """Outputs [M, 28] (float64) and the (Z, H) cache of the hidden layers."""
    layers = unpack(np.asarray(params, np.float32))
    h = np.asarray(x, np.float32)
    zs, hs = [], []
    for l, (W, b) in enumerate(layers):
        z = _op(h, bf16_ops) @ _op(W, bf16_ops) + b.astype(np.float64)
        if l < 3:
            zs.append(z)
            h = silu(z)
            if bf16_ops:
                h = h.astype(np.float32)  # the kernels keep f32 activations between products
            hs.append(h)
        else:
            return z, (zs, hs)


def loss_and_grad(params: np.ndarray, x: np.ndarray, t: np.ndarray, bf16_ops: bool = False):
    """MSE = mean((y - t)^2) and its gradient as a flat float64 vector."""
    layers = unpack(np.asarray(params, np.float32))
    y, (zs, hs) = forward(params, x, bf16_ops)
    M, O = y.shape
    d = y - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    dz = 2.0 * d / (M * O)
    if bf16_ops:
        dz = dz.astype(np.float32)
    ins = [np.asarray(x, np.float32)] + hs
    grads = [None] * 4
    for l in range(3, -1, -1):
        W = layers[l][0]
        gW = _op(ins[l], bf16_ops).T @ _op(dz, bf16_ops)
        gb = np.asarray(dz, np.float64).sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dh = _op(dz, bf16_ops) @ _op(W, bf16_ops).T
            dz = dh * silu_grad(zs[l - 1])
            if bf16_ops:
                dz = dz.astype(np.float32)
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat


def adam(params, grads, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32; returns (params, m, v)."""
    f = np.float32
    g = np.asarray(grads, f)
    m = f(b1) * np.asarray(m, f) + f(1.0 - b1) * g
    v = f(b2) * np.asarray(v, f) + f(1.0 - b2) * g * g
    c1 = f(1.0) - f(b1) ** f(step)
    c2 = f(1.0) - f(b2) ** f(step)
    w = np.asarray(params, f) - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return w.astype(f), m.astype(f), v.astype(f)


# ---------------------------------------------------------------------------
# The surrogate's input map (test infrastructure): one explicit step of the
# cell tick's rate laws (crates/marl-cell/src/cell.rs:223-329) from the
# pre-tick state, routed through the cell's wiring, in float32 with one IEEE
# operation per step in the order the HIP kernel uses.

F32 = np.float32
F32_EPS = F32(1.1920929e-07)


def explicit_step(I0, E, lt, rule, quiescent, prep_remaining, sim, gross=False):
    """(dI[16], dD[12], frel[16]): the explicit internal deltas, exchange fluxes
    and per-reaction flux / (substrate + 1e-6) of one cell.  I0: pre-tick
    internals [16], E: environment [12], lt: the voxel's light.  With gross,
    also (gI[16], gD[12]): each output's gross activity, the sum of the
    magnitudes of the terms that touch it."""
    dt = F32(sim.dt)
    dI = [F32(0)] * 16
    dD = [F32(0)] * 12
    gI = [F32(0)] * 16
    gD = [F32(0)] * 12
    frel = [F32(0)] * 16
    one = F32(1)
    for tp in rule["transport"]:
        e, k = int(tp["ext_species"]), int(tp["int_species"])
        if e >= 12 or k >= 16:
            continue
        ex, ik = F32(E[e]), F32(I0[k])
        up = F32(tp["uptake_rate"]) * ex / (one + ex)
        sec = F32(tp["secrete_rate"]) * ik / (one + ik)
        dI[k] = dI[k] + (up - sec) * dt
        dD[e] = dD[e] + (sec - up) * dt
        g = (abs(up) + abs(sec)) * dt
        gI[k] = gI[k] + g
        gD[e] = gD[e] + g

    def val(i):
        return F32(lt) if i == 15 else F32(I0[i])

    thr = F32(sim.active_reaction_threshold)
    active = 0
    for q, rx in enumerate(rule["reactions"]):
        v, si, pi, ci, co = F32(rx["v_max"]), int(rx["substrate"]), int(rx["product"]), int(rx["catalyst"]), \
            int(rx["cofactor"])
        if abs(v) > thr and si != pi:
            active += 1
        if abs(v) < thr or si == pi or si >= 16 or pi >= 16 or ci >= 16:
            continue
        sv, cv = val(si), val(ci)
        rate = v * (sv / (F32(rx["k_m"]) + sv + F32_EPS)) * (F32(sim.epsilon) + cv / (F32(rx["k_cat"]) + cv + F32_EPS))
        if co != 0xFF and co < 16:
            cof = val(co)
            rate = rate * (cof / (one + cof))
            dI[co] = dI[co] - F32(0.5) * rate * dt
            gI[co] = gI[co] + abs(F32(0.5) * rate * dt)
        flux = rate * dt
        dI[si] = dI[si] - flux
        dI[pi] = dI[pi] + flux
        gI[si] = gI[si] + abs(flux)
        gI[pi] = gI[pi] + abs(flux)
        frel[q] = flux / (sv + F32(1e-6))
    pm = one
    if prep_remaining > 0:
        evolved = max(F32(rule["division_prep_ticks"]), one)
        pm = F32(sim.prep_maintenance_multiplier) + max(F32(sim.base_division_prep) - evolved, F32(0)) * \
            F32(sim.rush_penalty_rate)
    decay = F32(sim.lambda_maintenance) * pm * dt * F32(I0[0])
    upkeep = F32(active) * F32(sim.reaction_maintenance) * dt
    dI[0] = dI[0] - decay - upkeep
    gI[0] = gI[0] + abs(decay) + abs(upkeep)
    if not quiescent:
        for ef in rule["effectors"]:
            k, e = int(ef["int_species"]), int(ef["ext_species"])
            if k >= 16 or e >= 12:
                continue
            ik = F32(I0[k])
            if ik > F32(ef["threshold"]):
                amount = F32(ef["rate"]) * ik / (one + ik) * dt
                dI[k] = dI[k] - amount
                dD[e] = dD[e] + amount
                gI[k] = gI[k] + abs(amount)
                gD[e] = gD[e] + abs(amount)
    out = (np.array(dI, F32), np.array(dD, F32), np.array(frel, F32))
    return out + (np.array(gI, F32), np.array(gD, F32)) if gross else out


def target_scales(gI, gD):
    """The 28 target scales of a cell: each output's gross activity plus 1e-5
    (f32), 0 for an output with no gross activity."""
    return np.array([F32(g) + F32(1e-5) if F32(g) > 0 else F32(0) for g in list(gI) + list(gD)], F32)


def scaled_targets(resid, scales):
    """Residuals over their scales, clipped to +-4, 0 where the scale is 0."""
    resid = np.asarray(resid, F32)
    out = np.zeros_like(resid)
    nz = scales > 0
    out[nz] = np.clip(resid[nz] / scales[nz], F32(-4), F32(4))
    return out
