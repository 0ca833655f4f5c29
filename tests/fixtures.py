"""Field and occupancy fixtures restated from the reference's own tests.

init_deterministic   crates/marl-field/src/field.rs:257-268
init_wave            crates/marl-gpu/tests/gpu_diffusion.rs:7-23
occupancy_*          crates/marl-gpu/tests/gpu_diffusion.rs:25-54
All arithmetic is f32 in the reference's order.
"""
import numpy as np

F = np.float32


def grid(dims):
    X, Y, Z = dims
    z, y, x, s = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), np.arange(12), indexing="ij")
    return x, y, z, s


def init_deterministic(dims):
    x, y, z, s = grid(dims)
    k = (x * 13 + y * 17 + z * 19 + s * 23) % 541
    return (k.astype(F) * F(0.001)).astype(F).reshape(-1)


def init_wave(dims):
    X = dims[0]
    x, y, z, s = grid(dims)
    wave = ((x * 17 + y * 31 + z * 43 + s * 59) % 997).astype(F) / F(997.0)
    boost = np.where((s == 7) & (x > X // 3) & (x < X // 3 + 16), F(2.0), F(0.0)).astype(F)
    return ((wave * (s.astype(F) + F(1.0))) * F(0.02) + boost).astype(F).reshape(-1)


def vidx(dims, x, y, z):
    return (z * dims[1] + y) * dims[0] + x


def occupancy_empty(dims):
    return np.zeros(dims[0] * dims[1] * dims[2], np.uint8)


def occupancy_center(dims):
    o = occupancy_empty(dims)
    o[vidx(dims, dims[0] // 2, dims[1] // 2, dims[2] // 2)] = 1
    return o


def occupancy_boundary_adjacent(dims):
    o = occupancy_empty(dims)
    o[vidx(dims, dims[0] // 2, dims[1] // 2, 1)] = 1
    return o


def occupancy_dense_cluster(dims):
    o = occupancy_empty(dims)
    cx, cy, cz = dims[0] // 2, dims[1] // 2, dims[2] // 2
    for z in range(cz - 1, cz + 2):
        for y in range(cy - 1, cy + 2):
            for x in range(cx - 1, cx + 2):
                o[vidx(dims, x, y, z)] = 1
    return o


# the six CPU/GPU equivalence cases of gpu_diffusion.rs:114-150: (name, occupancy, substeps)
REFERENCE_CASES = [
    ("empty occupancy", occupancy_empty, 2),
    ("center occupancy", occupancy_center, 2),
    ("boundary-adjacent occupancy", occupancy_boundary_adjacent, 2),
    ("dense cluster occupancy", occupancy_dense_cluster, 2),
    ("single substep ping-pong", occupancy_center, 1),
    ("default substeps", occupancy_empty, 10),
]
