"""GPU parity: the HIP path (through the C ABI and the pybind11 module) against
the CPU oracle on the same seeded inputs.

Tolerances: complex128 |d| <= 1e-10 * max(|ref|, 1e-3 * ||ref||_inf) (written
as ``close128``); complex64 within 1e-5 of the complex128 oracle (``close64``);
samples bit-exact when both sides sample the same amplitudes.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

N = pytest.importorskip("pennylane_lightning_gpu._native")


def close128(got, ref, rel=1e-10):
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.maximum(np.abs(ref), 1e-3 * max(np.max(np.abs(ref)), 1e-300))
    err = np.abs(got - ref)
    assert np.all(err <= rel * scale + 1e-300), f"max rel err {np.max(err / scale):.3e}"


def close64(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    assert np.max(np.abs(got - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref))), \
        f"max abs err {np.max(np.abs(got - ref)):.3e}"


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


GATE_WIRES = {
    "PauliX": 1, "PauliY": 1, "PauliZ": 1, "Hadamard": 1, "S": 1, "T": 1, "PhaseShift": 1, "RX": 1,
    "RY": 1, "RZ": 1, "Rot": 1, "CNOT": 2, "SWAP": 2, "CY": 2, "CZ": 2, "ControlledPhaseShift": 2,
    "CRX": 2, "CRY": 2, "CRZ": 2, "CRot": 2, "IsingXX": 2, "IsingYY": 2, "IsingZZ": 2,
    "SingleExcitation": 2, "SingleExcitationMinus": 2, "SingleExcitationPlus": 2, "CSWAP": 3,
    "Toffoli": 3, "DoubleExcitation": 4, "DoubleExcitationMinus": 4, "DoubleExcitationPlus": 4,
    "MultiRZ": 3,
}


def params_for(name, rng):
    return list(rng.uniform(-3, 3, N.PARAM_COUNT.get(name, 0)))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("name", sorted(GATE_WIRES))
def test_each_gate(name, dtype):
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    n = 9
    psi = rand_state(n, rng)
    nw = GATE_WIRES[name]
    for trial in range(4):
        # low, high and mixed target bits
        wires = [int(w) for w in rng.choice(n, nw, replace=False)]
        if trial == 0:
            wires = list(range(n - nw, n))       # lowest amplitude bits
        elif trial == 1:
            wires = list(range(nw))              # highest amplitude bits
        p = params_for(name, rng)
        for inv in (False, True):
            ops = [(name, wires, p, inv)]
            ref = po.apply_ops(n, psi, ops)
            s = N.State(n, dtype).upload(psi.astype(dtype))
            got = s.apply(ops).download()
            (close128 if dtype == np.complex128 else close64)(got, ref)


@pytest.mark.parametrize("n", [1, 2, 5, 12, 17])
def test_random_circuits(n):
    import test_oracle as T

    rng = np.random.default_rng(500 + n)
    ops = T.random_circuit(n, rng, depth=3) if n > 1 else [("RX", [0], [0.3], False), ("Hadamard", [0], [], False)]
    psi = rand_state(n, rng)
    ref = po.apply_ops(n, psi, ops)
    close128(N.State(n).upload(psi).apply(ops).download(), ref)
    close64(N.State(n, np.complex64).upload(psi.astype(np.complex64)).apply(ops).download(), ref)


def test_matrix_gates_all_sizes():
    rng = np.random.default_rng(8)
    n = 10
    psi = rand_state(n, rng)
    for k in (1, 2, 3, 4, 5, 6):
        A = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        Q, _ = np.linalg.qr(A)
        wires = [int(w) for w in rng.choice(n, k, replace=False)]
        for inv in (False, True):
            ops = [("QubitUnitary", wires, [], inv, Q)]
            close128(N.State(n).upload(psi).apply(ops).download(), po.apply_ops(n, psi, ops))


def test_multirz_wide_and_basis_init():
    rng = np.random.default_rng(9)
    n = 12
    psi = rand_state(n, rng)
    ops = [("MultiRZ", [0, 2, 3, 5, 7, 8, 11], [0.77], False)]
    close128(N.State(n).upload(psi).apply(ops).download(), po.apply_ops(n, psi, ops))
    s = N.State(n).set_basis(37)
    e = np.zeros(1 << n, complex)
    e[37] = 1
    assert np.array_equal(s.download(), e)
    s.reset()
    assert s.download()[0] == 1 and np.count_nonzero(s.download()) == 1
    # setStateVector permutation (Test_StateVectorCudaManaged_NonParam.cpp:934-1014)
    init = rand_state(3, rng)
    idx = [0, 2, 4, 6, 1, 3, 5, 7]
    vals = init[[1, 3, 5, 7, 0, 2, 4, 6]]
    expected = init.copy()
    for i in range(4):
        expected[[2 * i, 2 * i + 1]] = expected[[2 * i + 1, 2 * i]]
    np.testing.assert_array_equal(N.State(3).set_indexed(idx, vals).download(), expected)


def test_generators_match_oracle():
    rng = np.random.default_rng(10)
    n = 6
    psi = rand_state(n, rng)
    for name in ("RX", "RY", "RZ", "PhaseShift", "CRX", "CRY", "CRZ", "ControlledPhaseShift", "IsingXX",
                 "IsingYY", "IsingZZ", "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
                 "DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ"):
        nw = GATE_WIRES[name]
        wires = [int(w) for w in rng.choice(n, nw, replace=False)]
        s = N.State(n).upload(psi)
        scale = s.apply_generator(name, wires)
        close128(s.download(), po.apply_generator(n, psi, name, wires))
        assert scale == po.generator(name, nw)[1]


# ----------------------------------------------------------------------------- measurements

@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_expval_matrix(dtype):
    rng = np.random.default_rng(11)
    n = 11
    psi = rand_state(n, rng)
    s = N.State(n, dtype).upload(psi.astype(dtype))
    for k in (1, 2, 3, 4, 5, 7):
        A = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        H = A + A.conj().T
        wires = [int(w) for w in rng.choice(n, k, replace=False)]
        got = s.expval_matrix(wires, H)
        ref = po.expval_matrix(n, psi, wires, H)
        (close128 if dtype == np.complex128 else close64)(got.real, ref.real)


def test_expval_literal_fixture(literals):
    f = literals["dense_expval_8x8"]
    psi = np.array([complex(*a) for a in f["state"]])
    M = np.array([complex(*a) for a in f["matrix"]]).reshape(8, 8)
    e = N.State(3).upload(psi).expval_matrix(f["wires"], M)
    assert e.real == pytest.approx(f["expected"][0], rel=1e-7)
    assert e.imag == pytest.approx(f["expected"][1], rel=1e-7)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_pauli_words(dtype):
    rng = np.random.default_rng(12)
    n = 12
    psi = rand_state(n, rng)
    words, wires, coeffs = [], [], []
    for _ in range(40):
        k = int(rng.integers(1, 5))
        words.append("".join(rng.choice(list("XYZ"), k)))
        wires.append([int(x) for x in rng.choice(n, k, replace=False)])
        coeffs.append(rng.normal())
    words += ["ZZ", "ZZ", "XX"]  # shared flip masks
    wires += [[0, 11], [3, 4], [1, 10]]
    coeffs += [0.5, -0.25, 1.5]
    s = N.State(n, dtype).upload(psi.astype(dtype))
    tot, per = s.expval_pauli_words(words, wires, coeffs)
    rt, rp = po.expval_pauli_words(n, psi, words, wires, coeffs)
    (close128 if dtype == np.complex128 else close64)(per, rp)
    (close128 if dtype == np.complex128 else close64)(tot, rt)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [6, 11, 17])
def test_pauli_words_tiles_and_wide_words(dtype, n):
    """Tile passes (words grouped by their flip bits, every letter, phases
    i^nY of all four kinds, Z-only words) and the pair passes of words whose
    flips span more bits than a tile holds, on states smaller than, equal to
    and larger than one tile, against the oracle."""
    rng = np.random.default_rng(1000 + n)
    psi = rand_state(n, rng)
    words, wires = [], []
    for _ in range(70):
        k = int(rng.integers(1, min(n, 5) + 1))
        words.append("".join(rng.choice(list("XYZI"), k)))
        wires.append([int(x) for x in rng.choice(n, k, replace=False)])
    for ny in range(4):  # i^nY of every kind, flips on low and high bits
        words.append("Y" * (ny + 1) + "X")
        wires.append([int(x) for x in rng.choice(n, ny + 2, replace=False)])
    words += ["X" * n, "Y" * n, "XZ" * (n // 2), "Z" * n, ""]
    wires += [list(range(n)), list(range(n))[::-1], list(range(2 * (n // 2))), list(range(n)), []]
    coeffs = rng.normal(size=len(words)) + 0.3j * rng.normal(size=len(words))
    plan = N.pauli_plan(n, words, wires)
    if n == 17:
        assert plan["pair_launches"] >= 1 and plan["tile_launches"] >= 1
    s = N.State(n, dtype).upload(psi.astype(dtype))
    tot, per = s.expval_pauli_words(words, wires, coeffs)
    rt, rp = po.expval_pauli_words(n, psi, words, wires, coeffs)
    (close128 if dtype == np.complex128 else close64)(per, rp)
    (close128 if dtype == np.complex128 else close64)(tot, rt)


def tfim_csr(n, J=1.0, h=0.7):
    import test_oracle as T

    return T.tfim_csr(n, J, h)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_csr_expval_and_apply(literals, dtype):
    f = literals["csr_expval"]
    psi = np.array([complex(*a) for a in f["state"]])
    data = np.array([complex(*a) for a in f["data"]])
    c = N.CSR(f["indptr"], f["indices"], data, dtype=dtype)
    assert c.expval(N.State(3, dtype).upload(psi.astype(dtype))).real == pytest.approx(1.0, rel=1e-6)
    rng = np.random.default_rng(13)
    for n in (6, 14):
        H = tfim_csr(n)
        psi = rand_state(n, rng)
        s = N.State(n, dtype).upload(psi.astype(dtype))
        c = N.CSR(H.indptr, H.indices, H.data, dtype=dtype)
        ref = po.expval_csr(n, psi, H.indptr, H.indices.astype(np.int64), H.data)
        (close128 if dtype == np.complex128 else close64)(c.expval(s).real, ref.real)
        out = N.State(n, dtype)
        c.apply(out, s)
        (close128 if dtype == np.complex128 else close64)(out.download(), H @ psi)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_probs(dtype):
    rng = np.random.default_rng(14)
    n = 13
    psi = rand_state(n, rng)
    s = N.State(n, dtype).upload(psi.astype(dtype))
    for wires in ([0], [12], [3, 7], [7, 3], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9], list(range(n)),
                  [12, 0, 5, 6, 1, 2, 3, 4, 8, 9, 10, 11]):
        got = s.probs(wires)
        ref = po.probs(n, psi, wires)
        (close128 if dtype == np.complex128 else close64)(got, ref)


@pytest.mark.parametrize("n", [3, 10, 16])
def test_sampler_bit_exact(n):
    rng = np.random.default_rng(15 + n)
    psi = rand_state(n, rng)
    for shots, seed in ((1, 1), (1000, 7), (100000, 1337)):
        got = N.State(n).upload(psi).sample(shots, seed)
        ref = po.sample(n, psi, shots, seed)
        np.testing.assert_array_equal(got, ref)
    p64 = psi.astype(np.complex64)
    np.testing.assert_array_equal(N.State(n, np.complex64).upload(p64).sample(5000, 3), po.sample(n, p64, 5000, 3))
    bits = N.State(n).upload(psi).sample_bits(257, 9)
    idx = po.sample(n, psi, 257, 9)
    expect = (idx[:, None] >> (n - 1 - np.arange(n))[None, :].astype(np.uint64)) & np.uint64(1)
    np.testing.assert_array_equal(bits, expect)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_pauli_word_from(dtype):
    """dst = P src in one pass against the oracle's gate-by-gate P src, for
    words with every letter, low and high wires, and an odd number of Y."""
    rng = np.random.default_rng(61)
    n = 12
    psi = rand_state(n, rng)
    src = N.State(n, dtype).upload(psi.astype(dtype))
    for word, wires in (("Z", [0]), ("X", [n - 1]), ("Y", [5]), ("ZZ", [0, n - 1]), ("XYZI", [11, 3, 7, 0]),
                        ("YYY", [1, 10, 6]), ("XX", [2, 9]), ("", [])):
        dst = N.State(n, dtype).set_pauli_word_of(src, word, wires)
        ops = [({"X": "PauliX", "Y": "PauliY", "Z": "PauliZ"}[c], [w], [], False) for c, w in zip(word, wires)
               if c != "I"]
        ref = po.apply_ops(n, psi, ops)
        (close128 if dtype == np.complex128 else close64)(dst.download(), ref)
    with pytest.raises(N.LghError):
        src.set_pauli_word_of(src, "Z", [0])
    with pytest.raises(N.LghError):
        N.State(n, dtype).set_pauli_word_of(src, "ZZ", [3, 3])


def test_inner_axpy_scale():
    rng = np.random.default_rng(16)
    n = 10
    a, b = rand_state(n, rng), rand_state(n, rng)
    A, B = N.State(n).upload(a), N.State(n).upload(b)
    close128(A.inner(B), np.vdot(a, b))
    A.axpy(0.3 - 0.2j, B)
    close128(A.download(), a + (0.3 - 0.2j) * b)
    B.scale(1j)
    close128(B.download(), 1j * b)


# ----------------------------------------------------------------------------- adjoint

def layered_circuit(n, layers, rng, rot=("RX", "RY", "RZ"), ent="ring"):
    ops = []
    for _ in range(layers):
        for w in range(n):
            for g in rot:
                ops.append((g, [w], [float(rng.uniform(0, 2 * math.pi))], False))
        if ent == "ring":
            for w in range(n):
                ops.append(("CNOT", [w, (w + 1) % n], [], False))
        else:  # CZ brick
            for w in range(0, n - 1, 2):
                ops.append(("CZ", [w, w + 1], [], False))
            for w in range(1, n - 1, 2):
                ops.append(("CZ", [w, w + 1], [], False))
    return ops


def obs_states(n, psi, obs_list, dtype=np.complex128):
    out = []
    for factors in obs_list:
        v = psi
        for name, wires in factors:
            v = po.apply_ops(n, v, [(name, wires, [], False)])
        out.append(v)
    return out


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_adjoint_config1_vs_oracle(dtype):
    # 12 qubits, 4 layers RX,RY,RZ + CNOT ring, <Z_w> on all wires, 144 params
    n = 12
    rng = np.random.default_rng(1337)
    ops = layered_circuit(n, 4, rng)
    obs = [[("PauliZ", [w])] for w in range(n)]
    psi = po.apply_ops(n, po.zero_state(n), ops)
    bras = obs_states(n, psi, obs)
    tp = list(range(4 * n * 3))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    gpsi = N.State(n, dtype).upload(psi.astype(dtype))
    gb = [N.State(n, dtype).upload(b.astype(dtype)) for b in bras]
    got = N.adjoint_jacobian(gpsi, gb, ops, tp)
    (close128 if dtype == np.complex128 else close64)(got, ref)
    # forward state on the GPU from |0>
    s = N.State(n, dtype).reset().apply(ops)
    (close128 if dtype == np.complex128 else close64)(s.download(), psi)


def test_adjoint_all_generators_vs_oracle():
    n = 5
    rng = np.random.default_rng(21)
    init = rand_state(n, rng)
    gates = [("PhaseShift", [1]), ("CRX", [0, 2]), ("CRY", [3, 1]), ("CRZ", [2, 0]),
             ("ControlledPhaseShift", [1, 3]), ("IsingXX", [0, 4]), ("IsingYY", [2, 3]),
             ("IsingZZ", [3, 0]), ("SingleExcitation", [1, 2]), ("SingleExcitationMinus", [0, 3]),
             ("SingleExcitationPlus", [2, 1]), ("DoubleExcitation", [0, 1, 2, 3]),
             ("DoubleExcitationMinus", [4, 2, 1, 0]), ("DoubleExcitationPlus", [1, 3, 0, 2]),
             ("MultiRZ", [0, 2, 3]), ("RX", [2]), ("RY", [4]), ("RZ", [1])]
    ops = []
    for g, w in gates:
        ops.append((g, w, [float(rng.uniform(-2, 2))], bool(rng.integers(2))))
        ops.append(("Hadamard", [w[0]], [], False))
        ops.append(("Toffoli", [w[0], (w[0] + 1) % n, (w[0] + 2) % n], [], False))
    obs = [[("PauliX", [0]), ("PauliZ", [2])], [("PauliY", [3])], [("Hadamard", [4])]] * 4  # 12 bras: 2 batches
    psi = po.apply_ops(n, init, ops)
    bras = obs_states(n, psi, obs)
    tp = [0, 2, 3, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15, 16, 17]
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
    close128(got, ref)
    # apply_operations=True starts from the initial state
    got2 = N.adjoint_jacobian(N.State(n).upload(init), [N.State(n).upload(b) for b in bras], ops, tp,
                              apply_operations=True)
    close128(got2, ref)


def test_adjoint_rejects_multiparam():
    n = 2
    s = N.State(n).reset()
    b = N.State(n).reset()
    with pytest.raises(N.LghError):
        N.adjoint_jacobian(s, [b], [("Rot", [0], [0.1, 0.2, 0.3], False)], [0])


# ----------------------------------------------------------------------------- gate fusion

def mixed_circuit(n, rng, layers=3):
    ops = []
    one = ["RX", "RY", "RZ", "PhaseShift", "Hadamard", "S", "T", "PauliX", "PauliY", "PauliZ", "Rot"]
    two = ["CNOT", "CZ", "CY", "CRX", "CRY", "CRZ", "ControlledPhaseShift", "IsingXX", "IsingYY", "IsingZZ",
           "SWAP", "SingleExcitation", "CRot"]
    for _ in range(layers):
        for w in range(n):
            g = one[rng.integers(len(one))]
            ops.append((g, [w], list(rng.uniform(-3, 3, N.PARAM_COUNT.get(g, 0))), bool(rng.integers(2))))
        for _ in range(n):
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            g = two[rng.integers(len(two))]
            ops.append((g, [a, b], list(rng.uniform(-3, 3, N.PARAM_COUNT.get(g, 0))), bool(rng.integers(2))))
        a, b, c, d = (int(x) for x in rng.choice(n, 4, replace=False))
        ops.append(("Toffoli", [a, b, c], [], False))
        ops.append(("MultiRZ", [a, c, d], [float(rng.uniform(-3, 3))], False))
        ops.append(("DoubleExcitation", [a, b, c, d], [float(rng.uniform(-3, 3))], False))
        ops.append(("CSWAP", [d, a, b], [], False))
    return ops


@pytest.mark.parametrize("n", [10, 13, 16])
def test_fused_forward_matches_oracle(n):
    rng = np.random.default_rng(900 + n)
    ops = mixed_circuit(n, rng)
    psi = rand_state(n, rng)
    ref = po.apply_ops(n, psi, ops)
    try:
        for fuse in (True, False):
            N.set_fusion(fuse)
            close128(N.State(n).upload(psi).apply(ops).download(), ref)
            close64(N.State(n, np.complex64).upload(psi.astype(np.complex64)).apply(ops).download(), ref)
    finally:
        N.set_fusion(True)


@pytest.mark.parametrize("n", [11, 14])
def test_fused_adjoint_matches_oracle(n):
    rng = np.random.default_rng(950 + n)
    # differentiable gates carry at most one parameter
    ops = [o for o in mixed_circuit(n, rng, layers=2) if len(o[2]) <= 1]
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    obs = [[("PauliZ", [0])], [("PauliX", [n - 1]), ("PauliY", [3])], [("Hadamard", [5])]]
    bras = obs_states(n, psi, obs)
    npar = sum(1 for o in ops if o[2])
    tp = sorted(int(x) for x in rng.choice(npar, npar // 2, replace=False))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    try:
        for fuse in (True, False):
            N.set_fusion(fuse)
            got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
            close128(got, ref)
            got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                                       [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras],
                                       ops, tp)
            close64(got64, ref)
    finally:
        N.set_fusion(True)


def diagonal_heavy_circuit(n, rng, count):
    """Mostly diagonal gates (long diagonal runs inside the fused passes,
    with register, lane and other bits mixed) and a few RY that end runs."""
    one = ["RZ", "PhaseShift", "S", "T", "PauliZ", "RY"]
    two = ["CZ", "CRZ", "IsingZZ", "ControlledPhaseShift"]
    ops = []
    for _ in range(count):
        r = rng.random()
        if r < 0.45:
            g = one[rng.integers(len(one))]
            w = [int(rng.integers(n))]
        elif r < 0.9:
            g = two[rng.integers(len(two))]
            w = [int(x) for x in rng.choice(n, 2, replace=False)]
        else:
            g = "MultiRZ"
            w = [int(x) for x in rng.choice(n, 3, replace=False)]
        ops.append((g, w, list(rng.uniform(-3, 3, N.PARAM_COUNT.get(g, 0))), bool(rng.integers(2))))
    return ops


@pytest.mark.parametrize("bits", [(4, 3), (3, 3)])
@pytest.mark.parametrize("n", [12, 13])
def test_fused_diagonal_runs_match_oracle(n, bits):
    rng = np.random.default_rng(990 + n)
    ops = diagonal_heavy_circuit(n, rng, 240)
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    bras = obs_states(n, psi, [[("PauliZ", [0])], [("PauliX", [n - 1]), ("PauliY", [3])]])
    npar = sum(1 for o in ops if o[2])
    tp = list(range(0, npar, 2))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    try:
        N.set_tile_bits(*bits)
        close128(N.State(n).upload(init).apply(ops).download(), psi)
        close64(N.State(n, np.complex64).upload(init.astype(np.complex64)).apply(ops).download(), psi)
        got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
        close128(got, ref)
        got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                                   [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras], ops, tp)
        close64(got64, ref)
    finally:
        N.set_tile_bits(4, 0)


def controlled_fusable_circuit(n, rng, count):
    """Only gates that fit fused passes: rotations, controlled single-qubit
    gates and diagonal gates on overlapping wires.  Inside a pass, diagonal
    gates are moved past dense gates that commute with them (controls are
    used diagonally, targets are not), which these circuits exercise on every
    wire role."""
    one = ["RX", "RY", "RZ", "PhaseShift", "Hadamard", "S", "PauliX"]
    ctrl = ["CNOT", "CY", "CRX", "CRY"]
    diag = ["CZ", "CRZ", "ControlledPhaseShift", "IsingZZ"]
    ops = []
    for _ in range(count):
        r = rng.random()
        if r < 0.35:
            g, w = one[rng.integers(len(one))], [int(rng.integers(n))]
        elif r < 0.65:
            g, w = ctrl[rng.integers(len(ctrl))], [int(x) for x in rng.choice(n, 2, replace=False)]
        elif r < 0.95:
            g, w = diag[rng.integers(len(diag))], [int(x) for x in rng.choice(n, 2, replace=False)]
        else:
            g, w = "MultiRZ", [int(x) for x in rng.choice(n, 3, replace=False)]
        ops.append((g, w, list(rng.uniform(-3, 3, N.PARAM_COUNT.get(g, 0))), bool(rng.integers(2))))
    return ops


@pytest.mark.parametrize("bits", [(4, 3), (3, 3)])
@pytest.mark.parametrize("n", [12, 13])
def test_fused_reordered_passes_match_oracle(n, bits):
    rng = np.random.default_rng(1700 + n)
    ops = controlled_fusable_circuit(n, rng, 260)
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    bras = obs_states(n, psi, [[("PauliZ", [1])], [("PauliX", [n - 2]), ("PauliZ", [4])]])
    npar = sum(1 for o in ops if o[2])
    tp = list(range(npar))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    try:
        N.set_tile_bits(*bits)
        close128(N.State(n).upload(init).apply(ops).download(), psi)
        close64(N.State(n, np.complex64).upload(init.astype(np.complex64)).apply(ops).download(), psi)
        got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
        close128(got, ref)
        got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                                   [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras], ops, tp)
        close64(got64, ref)
    finally:
        N.set_tile_bits(4, 0)


@pytest.mark.parametrize("seed", range(6))
def test_fused_random_sweep_vs_oracle(seed):
    """Seeded random circuits of 14-17 qubits: mostly fusable gates (so the
    passes are long and their gates are regrouped), with a few gates that
    run on their own and split the passes.  Forward state and adjoint
    Jacobian against the oracle, complex128 and complex64."""
    rng = np.random.default_rng(4200 + seed)
    n = 14 + seed % 4
    ops = controlled_fusable_circuit(n, rng, 180 + 20 * seed)
    split = ["SWAP", "IsingXX", "IsingYY", "SingleExcitation"]
    for _ in range(6):
        g = split[rng.integers(len(split))]
        w = [int(x) for x in rng.choice(n, 2, replace=False)]
        ops.insert(int(rng.integers(len(ops))), (g, w, list(rng.uniform(-3, 3, N.PARAM_COUNT.get(g, 0))), False))
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    close128(N.State(n).upload(init).apply(ops).download(), psi)
    close64(N.State(n, np.complex64).upload(init.astype(np.complex64)).apply(ops).download(), psi)
    bras = obs_states(n, psi, [[("PauliZ", [0]), ("PauliZ", [n - 1])], [("PauliY", [n // 2])]])
    npar = sum(1 for o in ops if o[2])
    tp = sorted(int(x) for x in rng.choice(npar, (2 * npar) // 3, replace=False))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
    close128(got, ref)
    got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                               [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras], ops, tp)
    close64(got64, ref)


def test_fused_rotation_forms_special_angles():
    """Real rotations are stored as c [[1, -t], [t, 1]] or s [[t, -1], [1, t]]
    (whichever of cos, sin is larger) with the factors applied once per pass;
    the Jacobian factors are rescaled by the running product.  Angles at the
    boundaries of that choice: |cos| = |sin| (theta = pi/2), cos ~ 6e-17
    (theta = pi), the identity (theta = 0), negative cosines, and chains of
    rotations on one wire that the forward passes merge."""
    rng = np.random.default_rng(77)
    n = 12
    angles = [0.0, np.pi / 2, -np.pi / 2, np.pi, -np.pi, 3 * np.pi / 2, 2.5, -2.9, 4.0, 1e-9]
    ops = []
    for layer in range(3):
        for w in range(n):
            th = float(angles[(w + 3 * layer) % len(angles)])
            ops.append(("RY", [w], [th], bool((w + layer) % 3 == 0)))
            ops.append(("RZ", [w], [float(rng.uniform(-3, 3))], False))
            if w % 4 == 1:
                ops.append(("RY", [w], [float(angles[(w + layer + 5) % len(angles)])], False))
        for w in range(layer % 2, n - 1, 2):
            ops.append(("CZ", [w, w + 1], [], False))
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    close128(N.State(n).upload(init).apply(ops).download(), psi)
    close64(N.State(n, np.complex64).upload(init.astype(np.complex64)).apply(ops).download(), psi)
    bras = obs_states(n, psi, [[("PauliZ", [0]), ("PauliZ", [n - 1])], [("PauliX", [5])]])
    tp = list(range(sum(1 for o in ops if o[2])))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    close128(N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp), ref)
    got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                               [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras], ops, tp)
    close64(got64, ref)


def test_fused_adjoint_config2_shape():
    # RY/RZ + CZ brick (the 28-qubit benchmark circuit family) at 14 qubits
    from pennylane_lightning_gpu.workloads import ry_rz_cz_brick

    n = 14
    ops = ry_rz_cz_brick(n, 3, seed=5)
    psi = po.apply_ops(n, po.zero_state(n), ops)
    bra = po.apply_ops(n, psi, [("PauliZ", [0], [], False), ("PauliZ", [n - 1], [], False)])
    tp = list(range(sum(1 for o in ops if o[2])))
    ref = po.adjoint_jacobian(n, psi, [bra], ops, tp)
    got = N.adjoint_jacobian(N.State(n).reset().apply(ops), [N.State(n).upload(bra)], ops, tp)
    close128(got, ref)


@pytest.mark.parametrize("bits", [(3, 4), (4, 4), (3, 3)])
def test_fused_tile_bits_variants(bits):
    """Every tile height of the fused forward and adjoint kernels (the
    defaults are 4 forward, 3 adjoint) against the oracle."""
    rng = np.random.default_rng(977)
    n = 12
    ops = [o for o in mixed_circuit(n, rng, layers=2) if len(o[2]) <= 1]
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    bras = obs_states(n, psi, [[("PauliZ", [0])], [("PauliX", [n - 1]), ("PauliY", [3])]])
    npar = sum(1 for o in ops if o[2])
    tp = list(range(0, npar, 2))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    try:
        N.set_tile_bits(*bits)
        close128(N.State(n).upload(init).apply(ops).download(), psi)
        got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
        close128(got, ref)
        got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                                   [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras], ops, tp)
        close64(got64, ref)
    finally:
        N.set_tile_bits(4, 0)


@pytest.mark.parametrize("chunk", [2, 3, 8])
@pytest.mark.parametrize("n", [14, 17])
def test_fused_multitile_wavefronts(n, chunk):
    """Fused passes whose wavefronts step through `chunk` consecutive tiles
    (the launch shape of 26-30-qubit states), forced at 14 and 17 qubits by
    capping the grids' wavefronts; with chunk 3 the tile count is no multiple
    of the chunk, so the last wavefront takes a shorter run.  Forward state
    and adjoint Jacobian against the oracle, complex128 and complex64."""
    from pennylane_lightning_gpu.workloads import ry_rz_cz_brick

    rng = np.random.default_rng(5100 + 10 * n + chunk)
    ops = ry_rz_cz_brick(n, 3, seed=int(rng.integers(1 << 30))) + controlled_fusable_circuit(n, rng, 60)
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    bras = obs_states(n, psi, [[("PauliZ", [0]), ("PauliZ", [n - 1])], [("PauliX", [n // 2])]])
    npar = sum(1 for o in ops if o[2])
    tp = list(range(npar))
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    # tiles per pass: 2^(n - 6 - register bits); register bits 4 forward, 3 adjoint
    fwd_tiles, adj_tiles = 1 << (n - 10), 1 << (n - 9)
    cap_f, cap_a = -(-fwd_tiles // chunk), -(-adj_tiles // chunk)
    assert -(-fwd_tiles // cap_f) == chunk and -(-adj_tiles // cap_a) == chunk
    if chunk == 3:
        assert adj_tiles % chunk != 0
    try:
        N.set_pass_waves(cap_f, cap_a)
        close128(N.State(n).upload(init).apply(ops).download(), psi)
        close64(N.State(n, np.complex64).upload(init.astype(np.complex64)).apply(ops).download(), psi)
        got = N.adjoint_jacobian(N.State(n).upload(psi), [N.State(n).upload(b) for b in bras], ops, tp)
        close128(got, ref)
        got64 = N.adjoint_jacobian(N.State(n, np.complex64).upload(psi.astype(np.complex64)),
                                   [N.State(n, np.complex64).upload(b.astype(np.complex64)) for b in bras], ops, tp)
        close64(got64, ref)
    finally:
        N.set_pass_waves(0, 0)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_fused_deep_single_pass_rotation_scale(dtype):
    """A 10-qubit circuit of 40 layers of RY(~pi/2) + CZ brick is one forward
    pass holding 400 real rotations, each of which takes a factor of at most
    2^-1/2 out of the registers.  The running product is held above a floor
    (2^-20 for complex64, 2^-100 for complex128) by keeping later rotations
    whole; without it the registers overflow in complex64 and the closing
    factor underflows in complex128.  Forward state and a Jacobian over few
    trainable gates (so that the reverse passes are long as well) against
    the oracle."""
    n, layers = 10, 40
    rng = np.random.default_rng(4242)
    ops = []
    for _ in range(layers):
        for w in range(n):
            ops.append(("RY", [w], [float(np.pi / 2 + rng.uniform(-0.05, 0.05))], False))
        for w in range(0, n - 1, 2):
            ops.append(("CZ", [w, w + 1], [], False))
        for w in range(1, n - 1, 2):
            ops.append(("CZ", [w, w + 1], [], False))
    assert N.fusion_plan(n, ops, dtype=dtype)["passes"] == 1
    init = rand_state(n, rng)
    psi = po.apply_ops(n, init, ops)
    close = close128 if dtype == np.complex128 else close64
    close(N.State(n, dtype).upload(init.astype(dtype)).apply(ops).download(), psi)
    bras = obs_states(n, psi, [[("PauliZ", [0]), ("PauliZ", [n - 1])], [("PauliX", [3])]])
    tp = [0, 7, 150, 333, 399]
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    got = N.adjoint_jacobian(N.State(n, dtype).upload(psi.astype(dtype)),
                             [N.State(n, dtype).upload(b.astype(dtype)) for b in bras], ops, tp)
    assert np.all(np.isfinite(got))
    close(got, ref)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_autotuned_fused_passes_match_oracle(dtype):
    """lgh_autotune_fused times every candidate tile height and grid (and,
    for the adjoint, trainable gates per pass) on a 16-qubit circuit and
    keeps the fastest; forward states and adjoint Jacobians computed under
    the tuned choice still match the oracle."""
    n = 16
    rng = np.random.default_rng(1616)
    ops = []
    for _ in range(3):
        for w in range(n):
            ops.append(("RY", [w], [float(rng.uniform(-3, 3))], False))
            ops.append(("RZ", [w], [float(rng.uniform(-3, 3))], bool(w % 3 == 0)))
        for s0 in (0, 1):
            for w in range(s0, n - 1, 2):
                ops.append(("CZ", [w, w + 1], [], False))
    try:
        res = N.autotune_fused(n, ops, dtype=dtype)
        assert res["forward"]["tile_bits"] in (3, 4) and res["adjoint"]["tile_bits"] in (3, 4)
        assert res["forward"]["waves_per_cu"] in (512, 4096)
        assert res["adjoint"]["waves_per_cu"] in (128, 256, 512)
        assert res["adjoint"]["slots"] in (24, 48, 96)
        assert len(res["candidates"]) == 12
        assert {c["slots"] for c in res["candidates"] if c["dir"] == "adjoint"} == {24, 48, 96}
        assert res["forward"]["ms"] <= res["forward"]["default_ms"] + 1e-9
        assert res["adjoint"]["ms"] <= res["adjoint"]["default_ms"] + 1e-9
        psi = N.State(n, dtype).reset().apply(ops)
        ref = po.apply_ops(n, po.zero_state(n), ops)
        (close128 if dtype == np.complex128 else close64)(psi.download(), ref)
        bra = N.State(n, dtype)
        bra.copy_from(psi)
        bra.apply([("PauliZ", [0], [], False)])
        npar = sum(1 for o in ops if o[2])
        got = N.adjoint_jacobian(psi, [bra], ops, list(range(npar)))
        jref = po.adjoint_jacobian(n, ref, [po.apply_ops(n, ref, [("PauliZ", [0], [], False)])], ops,
                                   list(range(npar)))
        (close128 if dtype == np.complex128 else close64)(got, jref)
    finally:
        N.autotune_clear()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_adjoint_zwords_match_explicit_observable_states(dtype):
    """lgh_adjoint_jacobian_zwords forms Z_j psi inside the sweep's first
    reverse pass (or before a standalone first step, or up front without
    fusion); the Jacobian equals the one from explicit observable states
    Z_j psi bit for bit, and the oracle's within tolerance.  Cases: several
    observables (one the identity), multi-tile wavefront runs, a circuit that
    ends with a gate no fused pass takes (SWAP), and fusion switched off."""
    rng = np.random.default_rng(4242)
    n = 15
    ops = []
    for _ in range(3):
        for w in range(n):
            ops.append(("RY", [w], [float(rng.uniform(-3, 3))], False))
            ops.append(("RZ", [w], [float(rng.uniform(-3, 3))], bool(w % 4 == 1)))
        for s0 in (0, 1):
            for w in range(s0, n - 1, 2):
                ops.append(("CZ", [w, w + 1], [], False))
    zw = [[0, n - 1], [3], [], [2, 7, 11]]
    psi = N.State(n, dtype).reset().apply(ops)
    npar = sum(1 for o in ops if o[2])
    tp = list(range(0, npar, 3))

    def explicit(circ):
        bras = []
        for ws in zw:
            b = N.State(n, dtype)
            b.set_pauli_word_of(psi, "Z" * len(ws), ws)
            bras.append(b)
        return N.adjoint_jacobian(psi, bras, circ, tp)

    try:
        for label, circ, waves, fusion in (("fused", ops, 0, True), ("multi-tile", ops, 4, True),
                                           ("unfused", ops, 0, False),
                                           ("swap last", ops + [("SWAP", [1, 9], [], False)], 0, True)):
            N.set_fusion(fusion)
            N.set_pass_waves(0, waves)
            if label == "swap last":
                psi = N.State(n, dtype).reset().apply(circ)
            got = N.adjoint_jacobian_zwords(psi, zw, circ, tp)
            ref = explicit(circ)
            np.testing.assert_array_equal(got, ref, err_msg=label)
    finally:
        N.set_fusion(True)
        N.set_pass_waves(0, 0)
    # the oracle on the last circuit (its state is psi now)
    circ = ops + [("SWAP", [1, 9], [], False)]
    st = po.apply_ops(n, po.zero_state(n), circ)
    bras = [po.apply_ops(n, st, [("PauliZ", [w], [], False) for w in ws]) for ws in zw]
    jref = po.adjoint_jacobian(n, st, bras, circ, tp)
    (close128 if dtype == np.complex128 else close64)(got, jref)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_lean_passes_match_general_kernels(dtype):
    """Passes made only of uncontrolled real rotations and diagonal runs run
    in the lean kernels (lgh_set_lean_passes, on by default).  Forward state
    and adjoint Jacobian of the RY/RZ + CZ brick family (a lean circuit) and
    of a circuit with controlled gates (general kernels) match the oracle
    with the lean kernels on and off, and the two paths agree."""
    from pennylane_lightning_gpu.workloads import ry_rz_cz_brick

    close = close128 if dtype == np.complex128 else close64
    rng = np.random.default_rng(5150)
    for n, ops in ((14, ry_rz_cz_brick(14, 4, seed=77)), (13, controlled_fusable_circuit(13, rng, 150))):
        psi = po.apply_ops(n, po.zero_state(n), ops)
        bras = obs_states(n, psi, [[("PauliZ", [0]), ("PauliZ", [n - 1])]])
        npar = sum(1 for o in ops if o[2])
        tp = list(range(npar))
        ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
        got = {}
        try:
            for lean in (True, False):
                N.set_lean_passes(lean)
                st = N.State(n, dtype).reset().apply(ops)
                close(st.download(), psi)
                jac = N.adjoint_jacobian(st, [N.State(n, dtype).upload(b.astype(dtype)) for b in bras], ops, tp)
                close(jac, ref)
                got[lean] = np.asarray(jac)
        finally:
            N.set_lean_passes(True)
        close(got[True], got[False])
