"""Seeded synthetic circuits for the benchmark configurations.

Operations are tuples ``(name, wires, params, inverse)``.  Angles are drawn
with ``numpy.random.default_rng(seed).uniform(0, 2*pi)`` in operation order.
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 2 * math.pi


def layered_ring(n: int, layers: int, seed: int = 1337, rotations=("RX", "RY", "RZ")):
    """Per layer: each rotation on every wire, then a CNOT ring (w -> w+1 mod n).

    12 qubits x 4 layers is the plain CPU-runnable configuration
    (192 ops, 144 parameters, <Z_w> on every wire)."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(layers):
        for w in range(n):
            for g in rotations:
                ops.append((g, [w], [float(rng.uniform(0, TWO_PI))], False))
        if n > 1:
            for w in range(n):
                ops.append(("CNOT", [w, (w + 1) % n], [], False))
    return ops


def ry_rz_cz_brick(n: int, layers: int, seed: int = 1337):
    """Per layer: RY and RZ on every wire, then CZ on (0,1),(2,3),... and
    (1,2),(3,4),...  28 qubits x 6 layers gives 498 ops and 336 parameters."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(layers):
        for w in range(n):
            ops.append(("RY", [w], [float(rng.uniform(0, TWO_PI))], False))
            ops.append(("RZ", [w], [float(rng.uniform(0, TWO_PI))], False))
        for w in range(0, n - 1, 2):
            ops.append(("CZ", [w, w + 1], [], False))
        for w in range(1, n - 1, 2):
            ops.append(("CZ", [w, w + 1], [], False))
    return ops


def strongly_entangling(n: int, layers: int, seed: int = 1337):
    """PennyLane's StronglyEntanglingLayers as the device serializes it: per
    layer l, Rot(phi, theta, omega) on every wire, expanded to RZ(phi), RY(theta),
    RZ(omega) (reference _serialize.py:311-312), then CNOT(w, (w + r) mod n) for
    every wire with range r = (l mod (n-1)) + 1.  3 n L parameters; the
    reference's published adjoint benchmark differentiates <Z_w> on every wire
    of this circuit (.github/CHANGELOG.md:603-640)."""
    rng = np.random.default_rng(seed)
    ops = []
    for layer in range(layers):
        for w in range(n):
            for g in ("RZ", "RY", "RZ"):
                ops.append((g, [w], [float(rng.uniform(0, TWO_PI))], False))
        if n > 1:
            r = layer % (n - 1) + 1
            for w in range(n):
                ops.append(("CNOT", [w, (w + r) % n], [], False))
    return ops


def ladder(n: int, layers: int, seed: int = 1337):
    """Per layer: RX, RY, RZ on every wire, then CNOT(w, w+1) for w < n-1."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(layers):
        for w in range(n):
            for g in ("RX", "RY", "RZ"):
                ops.append((g, [w], [float(rng.uniform(0, TWO_PI))], False))
        for w in range(n - 1):
            ops.append(("CNOT", [w, w + 1], [], False))
    return ops


def random_pauli_hamiltonian(n: int, terms: int = 256, seed: int = 1337):
    """`terms` Pauli words of 1-4 distinct random wires, factors in {X,Y,Z},
    coefficients N(0,1)."""
    rng = np.random.default_rng(seed)
    words, wires, coeffs = [], [], []
    for _ in range(terms):
        k = int(rng.integers(1, 5))
        ws = [int(x) for x in rng.choice(n, k, replace=False)]
        words.append("".join(rng.choice(list("XYZ"), k)))
        wires.append(ws)
        coeffs.append(float(rng.normal()))
    return words, wires, np.array(coeffs)


def tfim_csr(n: int, J: float = 1.0, h: float = 0.7):
    """Open-chain transverse-field Ising H = -J sum Z_i Z_{i+1} - h sum X_i as
    0-based CSR arrays (indptr, indices, data) with wire 0 the most significant
    bit.  Rows are sorted by column; nnz = (n+1) 2^n."""
    N = 1 << n
    idx = np.arange(N, dtype=np.int64)
    shifts = (n - 1 - np.arange(n)).astype(np.int64)
    diag = np.zeros(N)
    for i in range(n - 1):
        bi = (idx >> shifts[i]) & 1
        bj = (idx >> shifts[i + 1]) & 1
        diag += -J * (1 - 2 * (bi ^ bj))
    cols = np.empty((N, n + 1), dtype=np.int64)
    vals = np.empty((N, n + 1), dtype=np.complex128)
    cols[:, 0] = idx
    vals[:, 0] = diag
    for w in range(n):
        cols[:, w + 1] = idx ^ (1 << int(shifts[w]))
        vals[:, w + 1] = -h
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, axis=1)
    vals = np.take_along_axis(vals, order, axis=1)
    indptr = np.arange(0, N * (n + 1) + 1, n + 1, dtype=np.int64)
    return indptr, cols.ravel(), vals.ravel()


def ring_template(n: int, layers: int = 3):
    """Gate structure of the batched workload's circuits: per layer RX, RY, RZ
    on every wire, then a CNOT ring.  Returns (ops, param_index) with one
    parameter per rotation: 12n gates and 9n parameters for 3 layers."""
    ops, pidx, k = [], [], 0
    for _ in range(layers):
        for w in range(n):
            for g in ("RX", "RY", "RZ"):
                ops.append((g, [w], [0.0], False))
                pidx.append(k)
                k += 1
        if n > 1:
            for w in range(n):
                ops.append(("CNOT", [w, (w + 1) % n], [], False))
                pidx.append(-1)
    return ops, pidx


def batch_widths(count: int = 4096, lo: int = 8, hi: int = 20, seed: int = 1337):
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.integers(lo, hi + 1, count)]
