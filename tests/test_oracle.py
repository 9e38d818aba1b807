"""Pins the CPU oracle (test infrastructure) before anything is checked against it.

Each check is chosen so that a plausible mistake (wrong sign, wire order,
transposed matrix, dropped term) fails: reference literals transcribed from
the reference's own tests (tests/golden/reference_literals.json), generators
checked through the matrix exponential, circuits checked against a brute-force
Kronecker-product unitary, the adjoint Jacobian against the exact
parameter-shift rule, Philox against its published known-answer vectors, and
the CSR path against the free-fermion ground energy of the open TFIM chain.
"""
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

from oracle import pyoracle as po

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1.0, -1.0]).astype(complex)
P0 = np.diag([1.0, 0.0]).astype(complex)
P1 = np.diag([0.0, 1.0]).astype(complex)
PAULI = {"I": I2, "X": X, "Y": Y, "Z": Z}


def kron(*ms):
    out = np.array([[1.0 + 0j]])
    for m in ms:
        out = np.kron(out, m)
    return out


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def embed(n, wires, M):
    """Full 2^n unitary for M acting on `wires` (first wire = MSB of M)."""
    k = len(wires)
    rest = [w for w in range(n) if w not in wires]
    perm = list(wires) + rest
    full = np.kron(M, np.eye(1 << (n - k)))
    # full acts on qubit order `perm`; permute back to 0..n-1
    T = full.reshape([2] * (2 * n))
    inv = np.argsort(perm)
    T = T.transpose(list(inv) + [n + i for i in inv])
    return T.reshape(1 << n, 1 << n)


# generator definitions from Pauli algebra: U(p) = expm(i * s * p * G)
GENS = {
    "RX": (1, X, -0.5),
    "RY": (1, Y, -0.5),
    "RZ": (1, Z, -0.5),
    "PhaseShift": (1, P1, 1.0),
    "CRX": (2, kron(P1, X), -0.5),
    "CRY": (2, kron(P1, Y), -0.5),
    "CRZ": (2, kron(P1, Z), -0.5),
    "ControlledPhaseShift": (2, kron(P1, P1), 1.0),
    "IsingXX": (2, kron(X, X), -0.5),
    "IsingYY": (2, kron(Y, Y), -0.5),
    "IsingZZ": (2, kron(Z, Z), -0.5),
    "SingleExcitation": (2, 0.5 * (kron(Y, X) - kron(X, Y)), -0.5),
    "SingleExcitationMinus": (2, 0.5 * (kron(Y, X) - kron(X, Y)) + kron(P0, P0) + kron(P1, P1), -0.5),
    "SingleExcitationPlus": (2, 0.5 * (kron(Y, X) - kron(X, Y)) - kron(P0, P0) - kron(P1, P1), -0.5),
    "MultiRZ": (3, kron(Z, Z, Z), -0.5),
}


def double_exc_gen(diag):
    # sigma_y between |0011> (3) and |1100> (12); `diag` on the 14 other states
    G = np.zeros((16, 16), dtype=complex)
    G[3, 12], G[12, 3] = -1j, 1j
    for i in range(16):
        if i not in (3, 12):
            G[i, i] = diag
    return G


GENS["DoubleExcitation"] = (4, double_exc_gen(0.0), -0.5)
GENS["DoubleExcitationMinus"] = (4, double_exc_gen(1.0), -0.5)
GENS["DoubleExcitationPlus"] = (4, double_exc_gen(-1.0), -0.5)


@pytest.mark.parametrize("name", sorted(GENS))
def test_gate_equals_exp_of_generator(name):
    nw, G, s = GENS[name]
    for p in (0.37, -1.3, 2.9):
        U = po.gate_matrix(name, nw, [p])
        np.testing.assert_allclose(U, scipy.linalg.expm(1j * s * p * G), atol=1e-13)
        Ud = po.gate_matrix(name, nw, [p], inverse=True)
        np.testing.assert_allclose(Ud @ U, np.eye(1 << nw), atol=1e-13)
    Go, so = po.generator(name, nw)
    assert so == s
    np.testing.assert_allclose(Go, G, atol=1e-15)


def test_double_excitation_action():
    # PennyLane definition: |0011> -> cos|0011> + sin|1100>
    p = 0.8
    U = po.gate_matrix("DoubleExcitation", 4, [p])
    e3 = np.zeros(16)
    e3[3] = 1
    out = U @ e3
    assert out[3] == pytest.approx(math.cos(p / 2)) and out[12] == pytest.approx(math.sin(p / 2))


FIXED = {
    "PauliX": X, "PauliY": Y, "PauliZ": Z,
    "Hadamard": np.array([[1, 1], [1, -1]]) / math.sqrt(2),
    "S": np.diag([1, 1j]), "T": np.diag([1, np.exp(1j * math.pi / 4)]),
    "CNOT": kron(P0, I2) + kron(P1, X), "CY": kron(P0, I2) + kron(P1, Y),
    "CZ": kron(P0, I2) + kron(P1, Z),
    "SWAP": (kron(I2, I2) + kron(X, X) + kron(Y, Y) + kron(Z, Z)) / 2,
    "Toffoli": kron(P0, I2, I2) + kron(P1, P0, I2) + kron(P1, P1, X),
    "CSWAP": kron(P0, I2, I2) + kron(P1, (kron(I2, I2) + kron(X, X) + kron(Y, Y) + kron(Z, Z)) / 2),
}


@pytest.mark.parametrize("name", sorted(FIXED))
def test_fixed_gates(name):
    M = FIXED[name]
    nw = int(round(math.log2(M.shape[0])))
    np.testing.assert_allclose(po.gate_matrix(name, nw), M, atol=1e-15)


def test_rot_and_crot():
    phi, th, om = 0.3, -1.1, 2.2
    ref = po.gate_matrix("RZ", 1, [om]) @ po.gate_matrix("RY", 1, [th]) @ po.gate_matrix("RZ", 1, [phi])
    np.testing.assert_allclose(po.gate_matrix("Rot", 1, [phi, th, om]), ref, atol=1e-14)
    np.testing.assert_allclose(po.gate_matrix("CRot", 2, [phi, th, om]),
                               kron(P0, I2) + kron(P1, ref), atol=1e-14)


def test_rx_reference_literals(literals):
    f = literals["rx_outputs"]
    for ang, exp, exp_adj in zip(f["angles"], f["expected"], f["expected_adj"]):
        out = po.apply_ops(1, po.zero_state(1), [("RX", [0], [ang], False)])
        np.testing.assert_allclose(out, [complex(*e) for e in exp], atol=1e-15)
        out = po.apply_ops(1, po.zero_state(1), [("RX", [0], [ang], True)])
        np.testing.assert_allclose(out, [complex(*e) for e in exp_adj], atol=1e-15)


def random_circuit(n, rng, depth=3):
    ops = []
    names1 = ["RX", "RY", "RZ", "PhaseShift", "Hadamard", "S", "T", "PauliY"]
    names2 = ["CNOT", "CZ", "CY", "SWAP", "CRX", "CRY", "CRZ", "IsingXX", "IsingYY", "IsingZZ",
              "ControlledPhaseShift", "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus"]
    for _ in range(depth):
        for w in range(n):
            nm = names1[rng.integers(len(names1))]
            ops.append((nm, [w], [rng.uniform(-3, 3)] if nm in ("RX", "RY", "RZ", "PhaseShift") else [],
                        bool(rng.integers(2))))
        for _ in range(n):
            a, b = rng.choice(n, 2, replace=False)
            nm = names2[rng.integers(len(names2))]
            fixed = nm in ("CNOT", "CZ", "CY", "SWAP")
            ops.append((nm, [int(a), int(b)], [] if fixed else [rng.uniform(-3, 3)], bool(rng.integers(2))))
        if n >= 3:
            a, b, c = rng.choice(n, 3, replace=False)
            ops.append(("Toffoli", [int(a), int(b), int(c)], [], False))
            ops.append(("CSWAP", [int(c), int(a), int(b)], [], False))
            ops.append(("MultiRZ", [int(a), int(c)], [rng.uniform(-3, 3)], False))
        if n >= 4:
            ws = [int(x) for x in rng.choice(n, 4, replace=False)]
            ops.append(("DoubleExcitation", ws, [rng.uniform(-3, 3)], False))
            ops.append(("CRot", ws[:2], list(rng.uniform(-3, 3, 3)), bool(rng.integers(2))))
    return ops


def brute_force(n, psi, ops):
    for name, wires, params, inv in ops:
        U = po.gate_matrix(name, len(wires), params, inv)
        psi = embed(n, wires, U) @ psi
    return psi


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_circuit_vs_kron_unitary(n):
    rng = np.random.default_rng(100 + n)
    ops = random_circuit(n, rng) if n > 1 else [("RX", [0], [0.3], False), ("Hadamard", [0], [], False)]
    psi = rand_state(n, rng)
    np.testing.assert_allclose(po.apply_ops(n, psi, ops), brute_force(n, psi, ops), atol=1e-12)


def test_wire_convention_basis_state():
    # PauliX on wire w of |0..0> sets amplitude index 2^(n-1-w)
    n = 4
    for w in range(n):
        out = po.apply_ops(n, po.zero_state(n), [("PauliX", [w], [], False)])
        assert abs(out[1 << (n - 1 - w)] - 1) < 1e-15


def test_matrix_op_and_multirz_large():
    rng = np.random.default_rng(5)
    n = 8
    psi = rand_state(n, rng)
    A = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    Q, _ = np.linalg.qr(A)
    out = po.apply_ops(n, psi, [("QubitUnitary", [5, 2], [], False, Q)])
    np.testing.assert_allclose(out, embed(n, [5, 2], Q) @ psi, atol=1e-12)
    out = po.apply_ops(n, psi, [("QubitUnitary", [5, 2], [], True, Q)])
    np.testing.assert_allclose(out, embed(n, [5, 2], Q.conj().T) @ psi, atol=1e-12)
    # MultiRZ on 7 wires goes through the parity path
    ws = [0, 1, 2, 3, 5, 6, 7]
    out = po.apply_ops(n, psi, [("MultiRZ", ws, [0.77], False)])
    G = kron(*[Z] * 7)
    np.testing.assert_allclose(out, embed(n, ws, scipy.linalg.expm(-0.5j * 0.77 * G)) @ psi, atol=1e-12)


# ----------------------------------------------------------------------------- measurements

def test_dense_expval_literal(literals):
    f = literals["dense_expval_8x8"]
    psi = np.array([complex(*a) for a in f["state"]])
    M = np.array([complex(*a) for a in f["matrix"]]).reshape(8, 8)
    e = po.expval_matrix(3, psi, f["wires"], M)
    assert e.real == pytest.approx(f["expected"][0], rel=1e-7)
    assert e.imag == pytest.approx(f["expected"][1], rel=1e-7)


def test_csr_expval_literal(literals):
    f = literals["csr_expval"]
    psi = np.array([complex(*a) for a in f["state"]])
    data = np.array([complex(*a) for a in f["data"]])
    e = po.expval_csr(3, psi, f["indptr"], f["indices"], data)
    assert e.real == pytest.approx(f["expected"], rel=1e-7)


def pauli_matrix(word):
    return kron(*[PAULI[c] for c in word])


def test_sparse_hamiltonian_literal(literals):
    f = literals["sparse_ham_3q"]
    psi = np.array([complex(*a) for a in f["state"]])
    H = sum(c * pauli_matrix(w) for c, w in f["terms"])
    Hc = sp.csr_matrix(H)
    e = po.expval_csr(3, psi, Hc.indptr, Hc.indices, Hc.data)
    assert e.real == pytest.approx(f["expected"], rel=1e-12)


def test_pauli_words_literals(literals):
    f = literals["sparse_pauli_words_2q"]
    ops = [tuple(o) for o in f["circuit"]]
    psi = po.apply_ops(2, po.zero_state(2), ops)
    for word, val in f["cases"]:
        tot, per = po.expval_pauli_words(2, psi, [word], [[0, 1]], [1.0])
        assert tot == pytest.approx(val, abs=1e-14)
        Hc = sp.csr_matrix(pauli_matrix(word))
        assert po.expval_csr(2, psi, Hc.indptr, Hc.indices, Hc.data).real == pytest.approx(val, abs=1e-14)
    g = literals["hamiltonian_2q"]
    for terms, val in g["cases"]:
        val = eval(val, {"cos": math.cos}) if isinstance(val, str) else val
        words = [t[0] for t in terms]
        tot, _ = po.expval_pauli_words(2, psi, words, [[0, 1]] * len(words), [t[1] for t in terms])
        assert tot == pytest.approx(val, abs=1e-14)


def test_pauli_words_vs_dense():
    rng = np.random.default_rng(7)
    n = 6
    psi = rand_state(n, rng)
    words, wires, coeffs = [], [], []
    for _ in range(20):
        k = int(rng.integers(1, 5))
        ws = [int(x) for x in rng.choice(n, k, replace=False)]
        words.append("".join(rng.choice(list("XYZ"), k)))
        wires.append(ws)
        coeffs.append(rng.normal())
    tot, per = po.expval_pauli_words(n, psi, words, wires, coeffs)
    dense = 0
    for w, ws, c, pw in zip(words, wires, coeffs, per):
        e = np.vdot(psi, embed(n, ws, pauli_matrix(w)) @ psi)
        assert pw == pytest.approx(e.real, abs=1e-13)
        dense += c * e.real
    assert tot == pytest.approx(dense, abs=1e-12)


def test_expval_closed_forms():
    # <Z0> = cos(t), <Z1> = cos(t)cos(p) after RX(t)0, RX(p)1, CNOT(0,1)
    # (tests/test_expval.py:62-74)
    t, p = 0.432, 0.123
    psi = po.apply_ops(3, po.zero_state(3), [("RX", [0], [t], False), ("RX", [1], [p], False),
                                             ("CNOT", [0, 1], [], False)])
    assert po.expval_matrix(3, psi, [0], Z).real == pytest.approx(math.cos(t))
    assert po.expval_matrix(3, psi, [1], Z).real == pytest.approx(math.cos(t) * math.cos(p))
    assert po.expval_matrix(3, psi, [1], Y).real == pytest.approx(-math.cos(t) * math.sin(p))


def test_probs_order_and_marginals():
    rng = np.random.default_rng(3)
    n = 5
    psi = rand_state(n, rng)
    full = np.abs(psi) ** 2
    # reversed-bit order: wires[0] is the LSB of the output index
    pr = po.probs(n, psi, [1, 3])
    T = full.reshape([2] * n)
    marg = T.sum(axis=(0, 2, 4))  # axes (w1, w3) with w1 as the first axis
    np.testing.assert_allclose(pr.reshape(2, 2), marg.T, atol=1e-15)
    assert po.probs(n, psi, list(range(n))).sum() == pytest.approx(1.0)
    # basis state |01000>: wire 1 set -> output bit 0 for wires [1, 3]
    b = np.zeros(1 << n, complex)
    b[1 << (n - 2)] = 1
    assert po.probs(n, b, [1, 3])[1] == 1.0


# ----------------------------------------------------------------------------- sampler

def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32_10
    assert list(po.philox_raw([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(po.philox_raw([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                        0x6D5451FD]
    assert list(po.philox_raw([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
    assert list(po.philox(0x299F31D0A4093822, 0x85A308D3243F6A88)) == \
        list(po.philox_raw([0x243F6A88, 0x85A308D3, 0, 0], [0xA4093822, 0x299F31D0]))


def test_sampler_distribution_and_sorted():
    rng = np.random.default_rng(11)
    n = 4
    psi = rand_state(n, rng)
    shots = 200000
    s = po.sample(n, psi, shots, seed=1234)
    assert np.all(np.diff(s.astype(np.int64)) >= 0)
    counts = np.bincount(s.astype(np.int64), minlength=16)
    p = np.abs(psi) ** 2
    chi2 = np.sum((counts - shots * p) ** 2 / (shots * p))
    assert chi2 < 50  # 15 dof: p-value ~ 1e-5
    # deterministic for a fixed seed; different for another
    assert np.array_equal(s, po.sample(n, psi, shots, seed=1234))
    assert not np.array_equal(s, po.sample(n, psi, shots, seed=99))
    # zero-probability outcomes never drawn
    b = np.zeros(16, complex)
    b[[3, 9]] = 1 / math.sqrt(2)
    s2 = po.sample(n, b, 1000, seed=5)
    assert set(np.unique(s2)) <= {3, 9}


# ----------------------------------------------------------------------------- adjoint

def obs_matrix(obs):
    """obs: list of (name, wires) factors of a tensor product."""
    return obs


def apply_obs(n, psi, factors):
    out = psi
    for name, wires in factors:
        out = po.apply_ops(n, out, [(name, wires, [], False)])
    return out


def adjoint(n, init, ops, obs_list, tp):
    psi = po.apply_ops(n, init, ops)
    bras = [apply_obs(n, psi, o) for o in obs_list]
    return po.adjoint_jacobian(n, psi, bras, ops, tp)


def test_adjoint_literals(literals):
    f = literals["adj_rx_z"]
    for p in f["params"]:
        j = adjoint(1, po.zero_state(1), [("RX", [0], [p], False)], [[("PauliZ", [0])]], [0])
        assert j[0, 0] == pytest.approx(-math.sin(p), abs=1e-14)
    f = literals["adj_ry_x"]
    for p in f["params"]:
        j = adjoint(1, po.zero_state(1), [("RY", [0], [p], False)], [[("PauliX", [0])]], [0])
        assert j[0, 0] == pytest.approx(math.cos(p), abs=1e-14)
    for key in ("adj_rx3_zzz", "adj_mixed_xxx", "adj_mixed_12op"):
        f = literals[key]
        ops = [tuple(o) for o in f["ops"]]
        obs = [[(o[0], o[1]) for o in f["obs"]]]
        j = adjoint(f["num_qubits"], po.zero_state(f["num_qubits"]), ops, obs, f["trainable"])
        np.testing.assert_allclose(j[0], f["expected"], atol=f["margin"])


def test_adjoint_rot_decomposed_literal(literals):
    f = literals["adj_rot_decomposed"]
    init = np.array([complex(*a) for a in f["init_state"]])
    thetas = np.linspace(*f["thetas_linspace"][:2], f["thetas_linspace"][2])
    for th, exp in zip(thetas, f["expected"]):
        ps = [th, th ** 3, math.sqrt(2) * th]
        ops = [("RZ", [0], [ps[0]], False), ("RY", [0], [ps[1]], False), ("RZ", [0], [ps[2]], False)]
        j = adjoint(1, init, ops, [[("PauliZ", [0])]], [0, 1, 2])
        np.testing.assert_allclose(j[0], exp, atol=f["margin"])


def test_adjoint_hamiltonian_literals(literals):
    for key in ("adj_ham_z0_z1", "adj_ham_3"):
        f = literals[key]
        n = f["num_qubits"]
        ops = [tuple(o) for o in f["ops"]]
        psi = po.apply_ops(n, po.zero_state(n), ops)
        bra = sum(c * apply_obs(n, psi, [tuple(t) for t in term]) for c, term in zip(f["coeffs"], f["terms"]))
        j = po.adjoint_jacobian(n, psi, [bra], ops, f["trainable"])
        p = [o[2][0] for o in ops]
        exp = eval(f["expected_fn"], {"sin": math.sin, "p0": p[0], "p2": p[-1]})
        np.testing.assert_allclose(j[0], np.atleast_1d(exp), atol=1e-14)


def expval_of(n, ops, init, obs_list):
    psi = po.apply_ops(n, init, ops)
    return np.array([np.vdot(psi, apply_obs(n, psi, o)).real for o in obs_list])


def test_adjoint_vs_parameter_shift_config1_shape():
    # 4-wire version of the 12-wire layered circuit: RX,RY,RZ per wire + CNOT ring
    n, layers = 4, 2
    rng = np.random.default_rng(1337)
    ops = []
    for _ in range(layers):
        for w in range(n):
            for g in ("RX", "RY", "RZ"):
                ops.append((g, [w], [rng.uniform(0, 2 * math.pi)], False))
        for w in range(n):
            ops.append(("CNOT", [w, (w + 1) % n], [], False))
    obs = [[("PauliZ", [w])] for w in range(n)]
    init = po.zero_state(n)
    npar = layers * n * 3
    j = adjoint(n, init, ops, obs, list(range(npar)))
    # exact two-term shift rule for exp(-i p/2 P) rotations
    pidx = [i for i, o in enumerate(ops) if o[2]]
    for t, oi in enumerate(pidx):
        plus = [list(o) for o in ops]
        minus = [list(o) for o in ops]
        plus[oi][2] = [ops[oi][2][0] + math.pi / 2]
        minus[oi][2] = [ops[oi][2][0] - math.pi / 2]
        d = (expval_of(n, [tuple(o) for o in plus], init, obs) -
             expval_of(n, [tuple(o) for o in minus], init, obs)) / 2
        np.testing.assert_allclose(j[:, t], d, atol=1e-12)


def test_adjoint_all_generators_vs_finite_difference():
    n = 4
    rng = np.random.default_rng(21)
    init = rand_state(n, rng)
    gates = [("PhaseShift", [1]), ("CRX", [0, 2]), ("CRY", [3, 1]), ("CRZ", [2, 0]),
             ("ControlledPhaseShift", [1, 3]), ("IsingXX", [0, 1]), ("IsingYY", [2, 3]),
             ("IsingZZ", [3, 0]), ("SingleExcitation", [1, 2]), ("SingleExcitationMinus", [0, 3]),
             ("SingleExcitationPlus", [2, 1]), ("DoubleExcitation", [0, 1, 2, 3]),
             ("DoubleExcitationMinus", [3, 2, 1, 0]), ("DoubleExcitationPlus", [1, 3, 0, 2]),
             ("MultiRZ", [0, 2, 3]), ("RX", [2])]
    ops = []
    for g, w in gates:
        ops.append((g, w, [rng.uniform(-2, 2)], bool(rng.integers(2))))
        ops.append(("Hadamard", [w[0]], [], False))
    obs = [[("PauliX", [0]), ("PauliZ", [2])], [("PauliY", [3])]]
    npar = len(gates)
    j = adjoint(n, init, ops, obs, list(range(npar)))
    h = 1e-6
    pidx = [i for i, o in enumerate(ops) if o[2]]
    for t, oi in enumerate(pidx):
        plus = [list(o) for o in ops]
        minus = [list(o) for o in ops]
        plus[oi][2] = [ops[oi][2][0] + h]
        minus[oi][2] = [ops[oi][2][0] - h]
        d = (expval_of(n, [tuple(o) for o in plus], init, obs) -
             expval_of(n, [tuple(o) for o in minus], init, obs)) / (2 * h)
        np.testing.assert_allclose(j[:, t], d, atol=1e-8)


def test_adjoint_trainable_subset_and_stateprep_skip():
    n = 2
    ops = [("RX", [0], [0.3], False), ("RY", [1], [0.5], False), ("CNOT", [0, 1], [], False),
           ("RZ", [1], [0.9], False)]
    obs = [[("PauliZ", [1])]]
    full = adjoint(n, po.zero_state(n), ops, obs, [0, 1, 2])
    sub = adjoint(n, po.zero_state(n), ops, obs, [0, 2])
    np.testing.assert_allclose(sub[0], full[0, [0, 2]], atol=1e-15)
    # a leading StatePrep op is skipped by the reverse sweep
    init = po.apply_ops(n, po.zero_state(n), [("Hadamard", [0], [], False)])
    ops2 = [("StatePrep", [0, 1], [], False, np.eye(4))] + ops
    psi = po.apply_ops(n, init, ops)
    bras = [apply_obs(n, psi, o) for o in obs]
    j1 = po.adjoint_jacobian(n, psi, bras, ops, [0, 1, 2])
    j2 = po.adjoint_jacobian(n, psi, bras, ops2, [0, 1, 2])
    np.testing.assert_allclose(j1, j2, atol=1e-15)


# ----------------------------------------------------------------------------- TFIM CSR

def tfim_csr(n, J=1.0, h=0.7):
    """Open-chain TFIM H = -J sum Z_i Z_{i+1} - h sum X_i as CSR (wire 0 = MSB)."""
    N = 1 << n
    idx = np.arange(N, dtype=np.int64)
    bits = (idx[:, None] >> (n - 1 - np.arange(n))) & 1
    zz = 1 - 2 * (bits[:, :-1] ^ bits[:, 1:])
    diag = -J * zz.sum(axis=1)
    rows = [idx]
    cols = [idx]
    vals = [diag.astype(complex)]
    for w in range(n):
        rows.append(idx)
        cols.append(idx ^ (1 << (n - 1 - w)))
        vals.append(np.full(N, -h, dtype=complex))
    H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    H.sort_indices()
    return H


def test_tfim_free_fermion_ground_energy():
    n, J, h = 8, 1.0, 0.7
    H = tfim_csr(n, J, h)
    # free-fermion closed form: E0 = -sum of singular values of the bidiagonal
    # matrix with h on the diagonal and J above it (open chain)
    B = np.diag(np.full(n, h)) + np.diag(np.full(n - 1, J), 1)
    E0 = -np.linalg.svd(B, compute_uv=False).sum()
    w, v = np.linalg.eigh(H.toarray())
    assert w[0] == pytest.approx(E0, abs=1e-10)
    psi0 = np.ascontiguousarray(v[:, 0].astype(complex))
    e = po.expval_csr(n, psi0, H.indptr, H.indices.astype(np.int64), H.data)
    assert e.real == pytest.approx(E0, abs=1e-10)
    assert abs(e.imag) < 1e-12
    # same operator as a sum of Pauli words
    words = ["ZZ"] * (n - 1) + ["X"] * n
    wires = [[i, i + 1] for i in range(n - 1)] + [[i] for i in range(n)]
    coeffs = [-J] * (n - 1) + [-h] * n
    tot, _ = po.expval_pauli_words(n, psi0, words, wires, coeffs)
    assert tot == pytest.approx(E0, abs=1e-10)
