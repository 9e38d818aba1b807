"""The pybind11 boundary (`lightning_gpu_qubit_ops`) on the GPU, exercised the
way the reference's own C++ and Python tests exercise theirs, with the
reference literals (tests/golden/reference_literals.json) and the oracle."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ops_mod = pytest.importorskip("pennylane_lightning_gpu.lightning_gpu_qubit_ops")

C128 = ops_mod.LightningGPU_C128
C64 = ops_mod.LightningGPU_C64


def state_of(sv, dtype=np.complex128):
    out = np.zeros(sv.dataLength(), dtype=dtype)
    sv.DeviceToHost(out, False)
    return out


def test_module_device_functions():
    assert ops_mod.DevPool.getTotalDevices() >= 1
    assert ops_mod.is_gpu_supported()
    assert ops_mod.get_gpu_arch() == (9, 5)
    assert ops_mod.get_gpu_arch_name() == "gfx950"
    t = ops_mod.DevTag(0)
    assert t.getDeviceID() == 0


def test_rx_literals(literals):
    f = literals["rx_outputs"]
    for ang, exp, exp_adj in zip(f["angles"], f["expected"], f["expected_adj"]):
        for adj, e in ((False, exp), (True, exp_adj)):
            sv = C128(1)
            sv.RX([0], adj, [ang])
            np.testing.assert_allclose(state_of(sv), [complex(*x) for x in e], atol=1e-15)
            sv = C128(1)
            sv.apply(["RX"], [[0]], [adj], [[ang]])
            np.testing.assert_allclose(state_of(sv), [complex(*x) for x in e], atol=1e-15)


def test_named_methods_match_oracle():
    rng = np.random.default_rng(3)
    n = 5
    v = rng.normal(size=32) + 1j * rng.normal(size=32)
    v /= np.linalg.norm(v)
    names = ["PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "SWAP", "CSWAP", "Toffoli", "CY", "CZ",
             "PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "Rot", "CRX", "CRY", "CRZ", "CRot", "IsingXX",
             "IsingYY", "IsingZZ", "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
             "DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ"]
    arity = {"CNOT": 2, "SWAP": 2, "CSWAP": 3, "Toffoli": 3, "CY": 2, "CZ": 2, "ControlledPhaseShift": 2,
             "CRX": 2, "CRY": 2, "CRZ": 2, "CRot": 2, "IsingXX": 2, "IsingYY": 2, "IsingZZ": 2,
             "SingleExcitation": 2, "SingleExcitationMinus": 2, "SingleExcitationPlus": 2, "DoubleExcitation": 4,
             "DoubleExcitationMinus": 4, "DoubleExcitationPlus": 4, "MultiRZ": 3}
    nparams = {"Rot": 3, "CRot": 3}
    for name in names:
        k = arity.get(name, 1)
        wires = [int(w) for w in rng.choice(n, k, replace=False)]
        npar = nparams.get(name, 0 if name in ("PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "SWAP",
                                                "CSWAP", "Toffoli", "CY", "CZ") else 1)
        p = list(rng.uniform(-3, 3, npar))
        for adj in (False, True):
            sv = C128(v)
            getattr(sv, name)(wires, adj, p)
            ref = po.apply_ops(n, v, [(name, wires, p, adj)])
            np.testing.assert_allclose(state_of(sv), ref, atol=1e-13)
    # explicit-matrix overload: apply(name, wires, inv, params, matrix)
    A = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    Q, _ = np.linalg.qr(A)
    sv = C128(v)
    sv.apply("QubitUnitary", [3, 1], False, [], Q.ravel())
    np.testing.assert_allclose(state_of(sv), po.apply_ops(n, v, [("QubitUnitary", [3, 1], [], False, Q)]),
                               atol=1e-13)
    with pytest.raises(ops_mod.PLException):
        C128(3).apply(["NotAGate"], [[0]], [False], [[]])


def test_expectation_overloads(literals):
    f = literals["dense_expval_8x8"]
    psi = np.array([complex(*a) for a in f["state"]])
    M = np.array([complex(*a) for a in f["matrix"]])
    sv = C128(3)
    sv.HostToDevice(psi, False)
    assert sv.ExpectationValue([0, 1, 2], M) == pytest.approx(1.263, rel=1e-7)
    # named observable with and without a matrix
    sv = C128(2)
    sv.RX([0], False, [0.4])
    sv.RY([1], False, [-0.2])
    assert sv.ExpectationValue("PauliZ", [0], [], np.array([], complex)) == pytest.approx(math.cos(0.4))
    Z = np.array([1, 0, 0, -1], complex)
    assert sv.ExpectationValue("PauliZ", [1], [], Z) == pytest.approx(math.cos(-0.2))
    # tensor observable via names list + kron matrix
    ZZ = np.kron(np.diag([1, -1]), np.diag([1, -1])).astype(complex).ravel()
    assert sv.ExpectationValue(["PauliZ", "PauliZ"], [0, 1], [[], []], ZZ) == pytest.approx(math.cos(0.4) * math.cos(-0.2))
    # Pauli words (tests/test_hamiltonian.py:41-63)
    g = literals["hamiltonian_2q"]
    for terms, val in g["cases"]:
        val = eval(val, {"cos": math.cos}) if isinstance(val, str) else val
        res = sv.ExpectationValue([t[0] for t in terms], [[0, 1]] * len(terms),
                                  np.array([t[1] for t in terms], complex))
        assert res == pytest.approx(val, abs=1e-12)
    # sparse literals (tests/test_hamiltonian_sparse.py:70-96)
    P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1, -1])}
    for word, val in literals["sparse_pauli_words_2q"]["cases"]:
        H = sp.csr_matrix(np.kron(P[word[0]], P[word[1]]).astype(complex))
        for cls, it in ((C128, np.int64), (C64, np.int32)):
            s = cls(2)
            s.RX([0], False, [0.4])
            s.RY([1], False, [-0.2])
            assert s.ExpectationValue(H.indptr.astype(it), H.indices.astype(it), H.data) == pytest.approx(val, abs=1e-6)


def test_probability_and_samples():
    sv = C128(3)
    sv.Hadamard([0], False, [])
    sv.RX([2], False, [0.7])
    p = sv.Probability([0, 2])
    ref = po.probs(3, state_of(sv), [0, 2])
    np.testing.assert_allclose(p, ref, atol=1e-15)
    s = sv.GenerateSamples(3, 1000)
    assert s.shape == (1000, 3) and set(np.unique(s)) <= {0, 1}
    assert np.all(s[:, 1] == 0)
    s1 = sv.GenerateSamples(3, 500, 42)
    s2 = sv.GenerateSamples(3, 500, 42)
    np.testing.assert_array_equal(s1, s2)


@pytest.mark.parametrize("cls", [C128, C64])
def test_generate_samples_rejects_other_widths(cls):
    """The (shots, num_wires) result only describes the state's samples when
    num_wires is its qubit count (Bindings.cpp:536-553); other widths raise
    instead of writing past the array."""
    sv = cls(4)
    sv.Hadamard([1], False, [])
    for bad in (0, 1, 3, 5, 64):
        with pytest.raises(ops_mod.PLException):
            sv.GenerateSamples(bad, 100)
        with pytest.raises(ops_mod.PLException):
            sv.GenerateSamples(bad, 100, 7)
    s = sv.GenerateSamples(4, 100, 7)
    assert s.shape == (100, 4) and np.all(s[:, [0, 2, 3]] == 0)


def obs(cls_prefix, kind, *args):
    return getattr(ops_mod, f"{kind}GPU_{cls_prefix}")(*args)


def run_adjoint(n, ops, observables, tp, init=None, prefix="C128"):
    sv = (C128 if prefix == "C128" else C64)(n)
    if init is not None:
        sv.HostToDevice(np.asarray(init, dtype=np.complex128 if prefix == "C128" else np.complex64), False)
    for name, wires, params, inv in ops:
        getattr(sv, name)(wires, inv, params)
    adj = getattr(ops_mod, f"AdjointJacobianGPU_{prefix}")()
    dt = np.float64 if prefix == "C128" else np.float32
    cdt = np.complex128 if prefix == "C128" else np.complex64
    opsl = adj.create_ops_list([o[0] for o in ops], [np.array(o[2], dtype=dt) for o in ops],
                               [o[1] for o in ops], [o[3] for o in ops], [np.array([], dtype=cdt) for _ in ops])
    return np.asarray(adj.adjoint_jacobian(sv, observables, opsl, tp))


@pytest.mark.parametrize("prefix", ["C128", "C64"])
def test_adjoint_reference_literals(literals, prefix):
    tol = 1e-7 if prefix == "C128" else 2e-5
    for p in literals["adj_rx_z"]["params"]:
        j = run_adjoint(1, [("RX", [0], [p], False)], [obs(prefix, "NamedObs", "PauliZ", [0])], [0], prefix=prefix)
        assert j[0, 0] == pytest.approx(-math.sin(p), abs=tol)
    for key in ("adj_rx3_zzz", "adj_mixed_xxx", "adj_mixed_12op"):
        f = literals[key]
        T = obs(prefix, "TensorProdObs", [obs(prefix, "NamedObs", o[0], o[1]) for o in f["obs"]])
        j = run_adjoint(f["num_qubits"], [tuple(o) for o in f["ops"]], [T], f["trainable"], prefix=prefix)
        np.testing.assert_allclose(j[0], f["expected"], atol=max(tol, f["margin"]))
    f = literals["adj_rot_decomposed"]
    init = [complex(*a) for a in f["init_state"]]
    for th, exp in zip(np.linspace(*f["thetas_linspace"][:2], f["thetas_linspace"][2]), f["expected"]):
        ps = [th, th ** 3, math.sqrt(2) * th]
        opsl = [("RZ", [0], [ps[0]], False), ("RY", [0], [ps[1]], False), ("RZ", [0], [ps[2]], False)]
        j = run_adjoint(1, opsl, [obs(prefix, "NamedObs", "PauliZ", [0])], [0, 1, 2], init=init, prefix=prefix)
        np.testing.assert_allclose(j[0], exp, atol=max(tol, 1e-7) * (1 if prefix == "C128" else 100))
    for key in ("adj_ham_z0_z1", "adj_ham_3"):
        f = literals[key]
        terms = [obs(prefix, "NamedObs", t[0][0], t[0][1]) for t in f["terms"]]
        dt = np.float64 if prefix == "C128" else np.float32
        H = obs(prefix, "Hamiltonian", np.array(f["coeffs"], dtype=dt), terms)
        opsl = [tuple(o) for o in f["ops"]]
        j = run_adjoint(f["num_qubits"], opsl, [H], f["trainable"], prefix=prefix)
        p = [o[2][0] for o in opsl]
        exp = eval(f["expected_fn"], {"sin": math.sin, "p0": p[0], "p2": p[-1]})
        np.testing.assert_allclose(j[0], np.atleast_1d(exp), atol=tol)


def test_adjoint_hermitian_equals_tensor():
    p = [-math.pi / 7, math.pi / 5, 2 * math.pi / 3]
    opsl = [("RX", [i], [p[i]], False) for i in range(3)]
    T = obs("C128", "TensorProdObs", [obs("C128", "NamedObs", "PauliZ", [0]), obs("C128", "NamedObs", "PauliZ", [1])])
    Hm = obs("C128", "HermitianObs", np.diag([1, -1, -1, 1]).astype(complex).ravel(), [0, 1])
    j1 = run_adjoint(3, opsl, [T], [0, 2])
    j2 = run_adjoint(3, opsl, [Hm], [0, 2])
    np.testing.assert_allclose(j1, j2, atol=1e-12)
    # sparse Hamiltonian observable in the adjoint
    Zd = np.diag([1.0, -1.0])
    H = sp.csr_matrix(np.kron(np.kron(Zd, Zd), np.eye(2)).astype(complex))
    S = obs("C128", "SparseHamiltonianObs".replace("Obs", ""), H.data, H.indices.astype(np.int64),
            H.indptr.astype(np.int64), [0, 1, 2])
    j3 = run_adjoint(3, opsl, [S], [0, 2])
    np.testing.assert_allclose(j1, j3, atol=1e-12)


def test_observable_repr_and_eq():
    a = obs("C128", "NamedObs", "PauliX", [3])
    b = obs("C128", "NamedObs", "PauliX", [3])
    c = obs("C128", "NamedObs", "RX", [1])
    assert repr(a) == "PauliX[3]"
    assert a == b and not (a == c)
    T = obs("C128", "TensorProdObs", [a, c])
    assert repr(T) == "PauliX[3] @ RX[1]"
    with pytest.raises(ops_mod.PLException):
        obs("C128", "TensorProdObs", [a, b])


def random_hermitian(dim, rng):
    A = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
    return A + A.conj().T


@pytest.mark.parametrize("prefix", ["C128", "C64"])
def test_adjoint_hermitian_and_sparse_vs_oracle(prefix):
    """Adjoint Jacobians of non-diagonal HermitianObs on 2 and 3 wires (given
    in an order other than ascending) and of a SparseHamiltonian over all
    wires against the oracle's adjoint with bra = O psi
    (ObservablesGPU.hpp:134-180, :469-587; AdjointDiffGPU.hpp:499-596)."""
    import scipy.sparse as sp

    rng = np.random.default_rng(3030)
    n = 6
    ops = []
    for layer in range(2):
        for w in range(n):
            ops.append(("RX", [w], [float(rng.uniform(-3, 3))], False))
            ops.append(("RY", [w], [float(rng.uniform(-3, 3))], bool(w % 2)))
        for w in range(n - 1):
            ops.append(("CNOT", [w, w + 1], [], False))
        ops.append(("CRZ", [layer, n - 1 - layer], [float(rng.uniform(-3, 3))], False))
        ops.append(("IsingXX", [1, 4], [float(rng.uniform(-3, 3))], False))
    psi = po.apply_ops(n, po.zero_state(n), ops)
    npar = sum(1 for o in ops if o[2])
    tp = list(range(0, npar, 2)) + [npar - 1]
    h2 = random_hermitian(4, rng)
    h3 = random_hermitian(8, rng)
    Hs = sp.random(1 << n, 1 << n, density=0.08, random_state=5, format="csr", dtype=np.float64)
    Hs = (Hs + 1j * sp.random(1 << n, 1 << n, density=0.08, random_state=6, format="csr")).tocsr()
    Hs = (Hs + Hs.conj().T).tocsr()
    Hs.sort_indices()
    cdt = np.complex128 if prefix == "C128" else np.complex64
    it = np.int64 if prefix == "C128" else np.int32
    observables = [
        obs(prefix, "HermitianObs", h2.astype(cdt).ravel(), [4, 1]),
        obs(prefix, "HermitianObs", h3.astype(cdt).ravel(), [2, 5, 0]),
        obs(prefix, "SparseHamiltonian", Hs.data.astype(cdt), Hs.indices.astype(it), Hs.indptr.astype(it),
            list(range(n))),
    ]
    bras = [po.apply_ops(n, psi, [("QubitUnitary", [4, 1], [], False, h2)]),
            po.apply_ops(n, psi, [("QubitUnitary", [2, 5, 0], [], False, h3)]),
            Hs @ psi]
    ref = po.adjoint_jacobian(n, psi, bras, ops, tp)
    got = run_adjoint(n, ops, observables, tp, prefix=prefix)
    if prefix == "C128":
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * max(1.0, np.max(np.abs(ref))))
    else:
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * max(1.0, np.max(np.abs(ref))))


@pytest.mark.parametrize("n,prefix", [(12, "C128"), (12, "C64"), (16, "C128")])
def test_adjoint_strongly_entangling_all_wires_vs_oracle(n, prefix):
    """The reference's published adjoint workload shape (.github/CHANGELOG.md:
    603-640): StronglyEntanglingLayers, Rot expanded to RZ RY RZ as the device
    serializes it (_serialize.py:311-312), <Z_w> on every wire; the full
    n x 3nL Jacobian against the oracle's adjoint."""
    from pennylane_lightning_gpu.workloads import strongly_entangling

    ops = strongly_entangling(n, 2, seed=77)
    psi = po.apply_ops(n, po.zero_state(n), ops)
    npar = 3 * n * 2
    bras = [po.apply_ops(n, psi, [("PauliZ", [w], [], False)]) for w in range(n)]
    ref = po.adjoint_jacobian(n, psi, bras, ops, list(range(npar)))
    got = run_adjoint(n, ops, [obs(prefix, "NamedObs", "PauliZ", [w]) for w in range(n)], list(range(npar)),
                      prefix=prefix)
    assert got.shape == (n, npar)
    tol = 1e-10 if prefix == "C128" else 1e-5
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol * max(1.0, np.max(np.abs(ref))))
