"""The LightningGPU device harness on the CPU: serialization, routing, the
trainable-parameter shift and the Hamiltonian row reduction, checked against
a stand-in backend that records what the harness asks of it (no GPU work).

The rules come from the reference device and serializer
(pennylane_lightning_gpu/lightning_gpu.py:519-960, _serialize.py:142-336).
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from pennylane_lightning_gpu import circuit as qc
from pennylane_lightning_gpu import _serialize as ser
from pennylane_lightning_gpu.lightning_gpu import LightningGPU, QuantumFunctionError, DeviceError

GATES = ["PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "CZ", "RX", "RY", "RZ", "Rot", "CRX",
         "PhaseShift", "IsingZZ", "Identity"]


class _Rec:
    """A constructed backend object: its class name and arguments."""

    def __init__(self, kind, *args):
        self.kind, self.args = kind, args

    def __repr__(self):
        return f"{self.kind}{self.args}"


def _obs_class(kind):
    return type(kind, (), {"__init__": lambda self, *a: _Rec.__init__(self, kind, *a),
                           "__repr__": _Rec.__repr__})


class FakeSV:
    """Records calls; holds no state."""

    def __init__(self, n_or_other):
        self.calls = []
        self.n = n_or_other if isinstance(n_or_other, int) else n_or_other.n

    def setBasisState(self, index, use_async):
        self.calls.append(("setBasisState", index))

    def resetGPU(self, use_async):
        self.calls.append(("resetGPU",))

    def setSeed(self, s):
        self.calls.append(("setSeed", s))

    def HostToDevice(self, arr, use_async):
        self.calls.append(("HostToDevice", np.array(arr)))

    def setStateVector(self, idx, state, use_async):
        self.calls.append(("setStateVector", np.array(idx), np.array(state)))

    def apply(self, name, wires, inv, params, mat):
        self.calls.append(("apply_matrix", name, list(wires), inv, np.array(mat)))

    def ExpectationValue(self, *args):
        self.calls.append(("ExpectationValue",) + args)
        return 0.25

    def Probability(self, wires):
        # p[i] with the backend's reversed bit order: index bit j <-> wires[j]
        return np.arange(1 << len(wires), dtype=float)


def _gate_method(name):
    def method(self, wires, inv, params):
        self.calls.append((name, list(wires), inv, list(params)))

    return method


for _g in GATES:  # named methods, as on the pybind11 class
    setattr(FakeSV, _g, _gate_method(_g))


class FakeAdj:
    def __init__(self):
        self.last = None

    def create_ops_list(self, names, params, wires, inverses, mats):
        return (names, params, wires, inverses, mats)

    def adjoint_jacobian(self, sv, obs, ops, tp):
        self.last = (obs, ops, list(tp))
        # row r, column c -> 100 r + tp[c]
        return np.array([[100.0 * r + t for t in tp] for r in range(len(obs))])

    adjoint_jacobian_batched = adjoint_jacobian


class FakeBackend:
    LightningGPU_C128 = FakeSV
    LightningGPU_C64 = FakeSV
    NamedObsGPU_C128 = _obs_class("Named")
    TensorProdObsGPU_C128 = _obs_class("Tensor")
    HamiltonianGPU_C128 = _obs_class("Ham")
    SparseHamiltonianGPU_C128 = _obs_class("Sparse")
    HermitianObsGPU_C128 = _obs_class("Herm")
    adj = FakeAdj()

    @classmethod
    def AdjointJacobianGPU_C128(cls):  # noqa: N802
        return cls.adj


def dev(wires=3, **kw):
    return LightningGPU(wires, _backend=FakeBackend, **kw)


def test_embed_matches_kronecker_permutation():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    # A on wires (c, a) of (a, b, c): brute force over basis states
    full = qc.embed(A, ["c", "a"], ["a", "b", "c"])
    ref = np.zeros((8, 8), dtype=complex)
    for i, j in itertools.product(range(8), range(8)):
        ia, ib, ic = (i >> 2) & 1, (i >> 1) & 1, i & 1
        ja, jb, jc = (j >> 2) & 1, (j >> 1) & 1, j & 1
        if ib == jb:
            ref[i, j] = A[(ic << 1) | ia, (jc << 1) | ja]
    np.testing.assert_allclose(full, ref)


def test_hamiltonian_matrix_and_wire_order():
    H = qc.Hamiltonian([0.5, -2.0], [qc.PauliZ(2) @ qc.PauliX(0), qc.PauliY(1)])
    assert H.wires == [2, 0, 1]
    Z, X, Y = qc.PauliZ(0).matrix(), qc.PauliX(0).matrix(), qc.PauliY(0).matrix()
    ref = 0.5 * np.kron(np.kron(Z, X), np.eye(2)) - 2.0 * np.kron(np.eye(4), Y)
    np.testing.assert_allclose(H.matrix(), ref)


def test_serialize_ops_rot_adjoint_matrix_and_stateprep():
    U = np.linalg.qr(np.random.default_rng(2).normal(size=(2, 2)) + 0j)[0]
    ops = [qc.StatePrep(np.array([1, 0, 0, 0]), [0, 1]), qc.Rot(0.1, 0.2, 0.3, 1),
           qc.Adjoint(qc.Operation("RX", [0], [0.4])), qc.Operation("QubitUnitary", [2], [], matrix=U),
           qc.Adjoint(qc.Operation("QubitUnitary", [2], [], matrix=U))]
    tape = qc.QuantumTape(ops, [qc.expval(qc.PauliZ(0))])
    (names, params, wires, inv, mats), use_sp = ser.TapeSerializer(FakeBackend, {0: 0, 1: 1, 2: 2}).operations(tape)
    assert use_sp
    assert names == ["RZ", "RY", "RZ", "RX", "QubitUnitary", "QubitUnitary"]
    assert [list(p) for p in params] == [[0.1], [0.2], [0.3], [0.4], [], []]
    assert inv == [False, False, False, True, False, False]  # matrices carry the adjoint themselves
    np.testing.assert_allclose(mats[4], U.ravel())
    np.testing.assert_allclose(mats[5], U.conj().T.ravel())
    assert wires == [[1], [1], [1], [0], [2], [2]]


def test_serialize_observables_offsets_split_and_identity():
    H = qc.Hamiltonian([1.0, 2.0, 3.0], [qc.PauliX(0), qc.PauliZ(1) @ qc.PauliY(2), qc.Identity(1)])
    tape = qc.QuantumTape([], [qc.expval(qc.PauliZ(0)), qc.expval(H), qc.expval(qc.Identity([1, 2]))])
    obs, offsets = ser.TapeSerializer(FakeBackend, {0: 0, 1: 1, 2: 2}).observables(tape)
    assert offsets == [0, 1, 4, 5]
    assert obs[0].kind == "Named" and obs[0].args == ("PauliZ", [0])
    assert [o.kind for o in obs[1:4]] == ["Ham"] * 3
    assert list(obs[2].args[0]) == [2.0] and obs[2].args[1][0].kind == "Tensor"
    assert obs[4].args == ("Identity", [1])  # Identity keeps its first wire only
    with pytest.raises(TypeError, match="Hermitian"):
        ser.TapeSerializer(FakeBackend, {0: 0}).rows(qc.Hermitian(np.eye(2), [0]))


def test_serialize_sparse_hamiltonian_csr_arrays():
    H = sp.random(8, 8, density=0.4, random_state=3, format="csr") + 0j
    o = ser.TapeSerializer(FakeBackend, {0: 0, 1: 1, 2: 2}).single(qc.SparseHamiltonian(H, [0, 1, 2]))
    data, idx, ptr, wires = o.args
    assert data.dtype == np.complex128 and idx.dtype == np.int64 and ptr.dtype == np.int64
    np.testing.assert_array_equal(ptr, H.indptr)
    assert wires == [0, 1, 2]


def test_apply_routes_named_inverse_matrix_and_identity():
    d = dev(["a", "b", "c"])
    U = np.eye(2)
    d.apply([qc.Operation("RX", ["b"], [0.5]), qc.Adjoint(qc.Operation("S", ["a"])), qc.Operation("Identity", ["c"]),
             qc.Operation("QubitUnitary", ["c"], [], matrix=U)])
    calls = [c for c in d._gpu_state.calls if c[0] not in ("setBasisState", "resetGPU")]
    assert calls[0] == ("RX", [1], False, [0.5])
    assert calls[1] == ("S", [0], True, [])
    assert calls[2][0] == "apply_matrix" and calls[2][1:4] == ("QubitUnitary", [2], False)
    assert len(calls) == 3


def test_state_preparation_rules():
    d = dev(3)
    d.apply([qc.BasisState([1, 0, 1], [0, 1, 2])])
    assert d._gpu_state.calls[-1] == ("setBasisState", 0b101)
    d.apply([qc.BasisState([1], [1])])
    assert d._gpu_state.calls[-1] == ("setBasisState", 0b010)
    v = np.array([0.6, 0.8j])
    d.apply([qc.StatePrep(v, [2])])
    name, idx, vals = d._gpu_state.calls[-1]
    assert name == "setStateVector" and list(idx) == [0, 1]
    d.apply([qc.StatePrep(v, [0])])
    assert list(d._gpu_state.calls[-1][1]) == [0, 4]
    with pytest.raises(DeviceError):
        d.apply([qc.Operation("PauliX", [0]), qc.BasisState([1], [0])])
    with pytest.raises(ValueError):
        d.apply([qc.StatePrep(np.array([1.0, 1.0]), [0])])


def test_expval_routing():
    d = dev(15)
    small = qc.Hamiltonian([1.0, 0.5], [qc.PauliX(0), qc.PauliZ(3) @ qc.PauliZ(4)])
    d.expval(small)
    call = d._gpu_state.calls[-1]
    assert len(call) == 3 and call[1] == [0, 3, 4] and call[2].size == 64  # dense matrix route
    big = qc.Hamiltonian([1.0, 0.5], [qc.PauliX(0) @ qc.PauliY(14), qc.PauliZ(13) @ qc.PauliZ(12)])
    big.wires = list(range(15))
    d.expval(big)
    call = d._gpu_state.calls[-1]
    assert call[1] == ["XY", "ZZ"] and call[2] == [[0, 14], [13, 12]]
    with pytest.raises(ValueError, match="Pauli word"):
        d.expval(qc.Hamiltonian([1.0] * 15, [qc.Hadamard(w) for w in range(15)]))
    d.expval(qc.PauliZ(2) @ qc.PauliX(5))
    call = d._gpu_state.calls[-1]
    assert call[1] == ["PauliZ", "PauliX"] and call[2] == [2, 5]
    d.expval(qc.SparseHamiltonian(sp.eye(4, format="csr"), [0, 1]))
    assert len(d._gpu_state.calls[-1]) == 4


def test_probability_transposes_reversed_bits_and_rejects_unsorted():
    d = dev(4)
    p = d.probability([0, 2, 3])
    # backend index bit j belongs to wires[j]; the device returns wires[0] as the MSB
    expect = [int(f"{i:03b}"[::-1], 2) for i in range(8)]
    np.testing.assert_array_equal(p, expect)
    with pytest.raises(RuntimeError, match="out-of-order"):
        d.probability([2, 0])


def test_adjoint_jacobian_tp_shift_reduction_and_shape():
    d = dev(3)
    H = qc.Hamiltonian([1.0, 2.0], [qc.PauliX(0), qc.PauliZ(1)])
    ops = [qc.StatePrep(np.array([1, 0]), [2]), qc.Operation("RX", [0], [0.1]), qc.Rot(0.1, 0.2, 0.3, 1),
           qc.Operation("CNOT", [0, 1]), qc.Operation("RY", [2], [0.5])]
    tape = qc.QuantumTape(ops, [qc.expval(qc.PauliZ(0)), qc.expval(H)], trainable_params=[0, 1, 3, 5])
    jac = d.adjoint_jacobian(tape)
    obs, _, tp = FakeBackend.adj.last
    # trainable 0 is the StatePrep array: no column; the rest shift by one
    assert tp == [0, 2, 4]
    assert len(obs) == 3  # Z0 and the two split terms of H
    # rows: 100 r + tp; H's rows (1 and 2) are summed to 300 + 2 tp; the StatePrep column stays 0
    assert np.allclose(np.array(jac), [[0, 0, 2, 4], [0, 300, 304, 308]])


def test_adjoint_checks():
    d = dev(2)
    t = qc.QuantumTape([qc.Operation("RX", [0], [0.1])], [qc.var(qc.PauliZ(0))])
    with pytest.raises(QuantumFunctionError, match="mixed"):
        d.adjoint_jacobian(t)
    t = qc.QuantumTape([qc.Operation("RX", [0], [0.1])], [qc.expval(qc.Hermitian(np.eye(2), [0]))])
    with pytest.raises(QuantumFunctionError, match="Hermitian"):
        d.adjoint_jacobian(t)
    t = qc.QuantumTape([qc.Operation("CRot", [0, 1], [0.1, 0.2, 0.3])], [qc.expval(qc.PauliZ(0))])
    with pytest.raises(QuantumFunctionError, match="CRot"):
        d.adjoint_jacobian(t)
    t = qc.QuantumTape([qc.Operation("RX", [0], [0.1])], [qc.expval(qc.PauliZ(0))], trainable_params=[])
    assert np.array(d.adjoint_jacobian(t)).shape == ()


def test_vjp_builds_weighted_observable():
    d = dev(2)
    ms = [qc.expval(qc.PauliZ(0)), qc.expval(qc.PauliX(1))]
    tape = qc.QuantumTape([qc.Operation("RX", [0], [0.1]), qc.Operation("RY", [1], [0.2])], ms)
    fn = d.vjp(ms, np.array([0.5, -1.5]))
    out = fn(tape)
    obs, _, tp = FakeBackend.adj.last
    assert [o.kind for o in obs] == ["Ham", "Ham"] and [list(o.args[0]) for o in obs] == [[0.5], [-1.5]]
    assert np.allclose(np.array(out), [100, 102])  # rows 0 and 1 summed: (0+100, 1+101)
    assert np.allclose(d.vjp(ms, np.zeros(2))(tape), 0)
    with pytest.raises(ValueError):
        d.vjp(ms, np.array([1.0]))
