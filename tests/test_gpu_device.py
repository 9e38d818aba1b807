"""The LightningGPU device harness on the GPU against the oracle: state
preparation, gates, every expectation route, variance, probabilities,
samples, the adjoint Jacobian (split Hamiltonians, a leading state
preparation, Rot, observable batching) and the VJP.

Tolerances as in test_gpu_parity.py: complex128 results within
1e-10 * max(|ref|, 1e-3 ||ref||_inf) of the oracle (``close128``).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pyoracle as po
from pennylane_lightning_gpu import circuit as qc

pytestmark = pytest.mark.gpu

pytest.importorskip("pennylane_lightning_gpu.lightning_gpu_qubit_ops")
from pennylane_lightning_gpu.lightning_gpu import LightningGPU  # noqa: E402

from _tol import close128  # noqa: E402


def close_scalar(got, ref, tol=1e-10):
    assert abs(got - ref) <= tol * max(1.0, abs(ref)), f"{got} vs {ref}"


def rand_state(k, rng):
    v = rng.normal(size=1 << k) + 1j * rng.normal(size=1 << k)
    return v / np.linalg.norm(v)


def to_oracle(ops, wire_map):
    """circuit ops -> oracle tuples (Rot expanded, adjoints as inverse flags)."""
    out = []
    for o in ops:
        if o.name in ("StatePrep", "BasisState", "Identity"):
            continue
        for s in (o.expand().operations if o.name == "Rot" else [o]):
            inv = hasattr(s, "base")
            base = s.base if inv else s
            w = [wire_map[x] for x in base.wires]
            if base.name == "QubitUnitary":
                out.append((base.name, w, [], False, s.matrix()))
            else:
                out.append((base.name, w, list(base.parameters), inv))
    return out


def pl_probs(n, psi, wires):
    """Marginal probabilities of increasing device wires, first wire most
    significant (PennyLane order), straight from |psi|^2."""
    p = np.abs(np.asarray(psi).reshape([2] * n)) ** 2
    return p.sum(axis=tuple(i for i in range(n) if i not in wires)).reshape(-1)


def full_matrix(obs, wires):
    return qc.embed(obs.matrix(), obs.wires, wires)


def circuit_ops(labels, rng):
    a, b, c, d, e, f = labels
    U = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    return [
        qc.StatePrep(rand_state(2, rng), [c, a]),
        qc.Operation("Hadamard", [b]), qc.Operation("RX", [d], [0.3]), qc.Rot(0.2, -0.7, 1.1, e),
        qc.Operation("CNOT", [b, f]), qc.Adjoint(qc.Operation("RY", [a], [0.9])), qc.Operation("CRZ", [f, c], [0.4]),
        qc.Operation("QubitUnitary", [e, b], [], matrix=U), qc.Operation("IsingZZ", [a, d], [-0.6]),
        qc.Operation("PhaseShift", [f], [0.8]), qc.Operation("Identity", [c]), qc.Operation("RZ", [b], [1.3]),
    ]


def reference_state(labels, ops):
    wm = {w: i for i, w in enumerate(labels)}
    prep = ops[0]
    n = len(labels)
    psi = np.zeros(1 << n, dtype=complex)
    sub = [wm[w] for w in prep.wires]
    for j, amp in enumerate(np.asarray(prep.parameters[0])):
        idx = 0
        for k, w in enumerate(sub):
            idx |= ((j >> (len(sub) - 1 - k)) & 1) << (n - 1 - w)
        psi[idx] = amp
    return po.apply_ops(n, psi, to_oracle(ops, wm)), wm


LABELS = ["q0", "q1", "anc", 7, "x", "y"]


def test_execute_state_and_measurements():
    rng = np.random.default_rng(11)
    ops = circuit_ops(LABELS, rng)
    ref, wm = reference_state(LABELS, ops)
    dev = LightningGPU(LABELS)
    dev.apply(ops)
    close128(dev.state, ref)
    n = len(LABELS)
    H = qc.Hermitian(np.array([[1.0, 0.5 - 0.2j], [0.5 + 0.2j, -0.3]]), ["x"])
    obs = [qc.PauliZ("q0"), qc.PauliX(7) @ qc.PauliY("anc"), qc.Hadamard("y"), qc.Identity("q1"), H,
           qc.Projector([1, 0], ["q1", "y"]),
           qc.Hamiltonian([0.4, -1.3, 2.0], [qc.PauliZ("q0") @ qc.PauliZ("y"), qc.PauliX("anc"), qc.Identity(7)])]
    for o in obs:
        M = full_matrix(o, LABELS)
        exact = float(np.real(np.vdot(ref, M @ ref)))
        close_scalar(dev.expval(o), exact)
        var = float(np.real(np.vdot(ref, M @ (M @ ref)))) - exact ** 2
        close_scalar(dev.var(o), var)
    # sparse Hamiltonian over all wires
    A = sp.random(1 << n, 1 << n, density=0.05, random_state=4, format="csr")
    Hs = (A + A.T).tocsr().astype(np.complex128)
    close_scalar(dev.expval(qc.SparseHamiltonian(Hs, LABELS)), float(np.real(np.vdot(ref, Hs @ ref))))
    # probabilities on increasing device wires, PennyLane order
    for ws in (["q0", "anc", "y"], [7], LABELS):
        close128(dev.probability(ws), pl_probs(n, ref, [wm[w] for w in ws]))


def test_hamiltonian_pauli_word_route_15_wires():
    rng = np.random.default_rng(5)
    n = 15
    psi = rand_state(n, rng)
    dev = LightningGPU(n)
    dev.syncH2D(psi)
    terms = [qc.PauliX(0) @ qc.PauliY(14), qc.PauliZ(3) @ qc.PauliZ(9) @ qc.PauliX(2), qc.PauliY(7), qc.Identity(1)]
    coeffs = [0.7, -1.1, 0.25, 2.0]
    H = qc.Hamiltonian(coeffs, terms)
    H.wires = list(range(n))  # spans all 15 wires: the Pauli-word route
    words = ["XY", "ZZX", "Y", "I"]
    wires = [[0, 14], [3, 9, 2], [7], [1]]
    close_scalar(dev.expval(H), po.expval_pauli_words(n, psi, words, wires, coeffs)[0])


@pytest.mark.parametrize("batch_obs", [False, True, 2])
def test_adjoint_jacobian_matches_oracle(batch_obs):
    rng = np.random.default_rng(21)
    ops = circuit_ops(LABELS, rng)
    ref, wm = reference_state(LABELS, ops)
    n = len(LABELS)
    H = qc.Hamiltonian([0.4, -1.3, 0.8], [qc.PauliZ("q0") @ qc.PauliZ("y"), qc.PauliX("anc"), qc.PauliY(7)])
    meas = [qc.expval(qc.PauliZ("q1")), qc.expval(H), qc.expval(qc.PauliX("x") @ qc.PauliZ("q0"))]
    tape = qc.QuantumTape(ops, meas)  # every parameter trainable, the StatePrep array included
    dev = LightningGPU(LABELS, batch_obs=batch_obs)
    jac = np.array(dev.adjoint_jacobian(tape))
    oops = to_oracle(ops, wm)
    npar = sum(1 for o in oops if len(o) == 4 and o[2])
    bras = [full_matrix(m.obs, LABELS) @ ref for m in meas]
    expect = po.adjoint_jacobian(n, ref, bras, oops, list(range(npar)))
    assert jac.shape == (3, npar + 1)
    np.testing.assert_array_equal(jac[:, 0], 0.0)  # the StatePrep parameter
    close128(jac[:, 1:], expect)
    # a subset of trainable parameters, through the device state
    tape2 = qc.QuantumTape(ops, meas, trainable_params=[2, 5, 6])
    jac2 = np.array(dev.adjoint_jacobian(tape2, use_device_state=True))
    close128(jac2, expect[:, [1, 4, 5]])


def test_vjp_is_dy_dot_jacobian():
    rng = np.random.default_rng(31)
    ops = circuit_ops(LABELS, rng)
    meas = [qc.expval(qc.PauliZ("q1")), qc.expval(qc.PauliY("y") @ qc.PauliX(7))]
    tape = qc.QuantumTape(ops, meas, trainable_params=[1, 2, 3, 4, 5, 6, 7, 8])
    dev = LightningGPU(LABELS)
    jac = np.array(dev.adjoint_jacobian(tape))
    dy = np.array([0.3, -1.7])
    close128(np.array(dev.vjp(meas, dy)(tape)), dy @ jac)


def test_samples_estimate_expvals_and_probabilities():
    rng = np.random.default_rng(41)
    ops = circuit_ops(LABELS, rng)
    ref, wm = reference_state(LABELS, ops)
    shots = 200000
    dev = LightningGPU(LABELS, shots=shots, seed=7)
    dev.apply(ops)
    for o in (qc.PauliZ("q0"), qc.PauliX(7) @ qc.PauliY("anc"), qc.Hadamard("y")):
        M = full_matrix(o, LABELS)
        exact = float(np.real(np.vdot(ref, M @ ref)))
        assert abs(dev.expval(o) - exact) < 6 / math.sqrt(shots)
    close128(dev.state, ref)  # sampling rotated a copy, not the device state
    p = dev.probability(["q1", "x"])
    assert np.max(np.abs(p - pl_probs(len(LABELS), ref, [wm["q1"], wm["x"]]))) < 6 / math.sqrt(shots)
