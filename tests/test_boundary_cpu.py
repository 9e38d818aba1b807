"""CPU-side checks of the drop-in boundary (no GPU calls):
the C-ABI library loads and exports every function declared in
include/lgh.h, the host-side gate compiler produces the oracle's matrices,
the pybind11 module has the reference's names, and errors map to PLException.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lgh.h")

N = pytest.importorskip("pennylane_lightning_gpu._native")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lgh_[a-z0-9_]+)\s*\(", text)))


def test_header_functions_exported():
    names = declared_functions()
    assert len(names) >= 35
    lib = N.lib()
    missing = [f for f in names if not hasattr(lib, f)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (lgh_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_abi_version_and_no_device_here():
    assert N.lib().lgh_abi_version() == 1
    assert N.device_count() >= 0


GATE_ARITY = {
    "PauliX": 1, "PauliY": 1, "PauliZ": 1, "Hadamard": 1, "S": 1, "T": 1, "PhaseShift": 1, "RX": 1,
    "RY": 1, "RZ": 1, "Rot": 1, "CNOT": 2, "SWAP": 2, "CY": 2, "CZ": 2, "ControlledPhaseShift": 2,
    "CRX": 2, "CRY": 2, "CRZ": 2, "CRot": 2, "IsingXX": 2, "IsingYY": 2, "IsingZZ": 2,
    "SingleExcitation": 2, "SingleExcitationMinus": 2, "SingleExcitationPlus": 2, "CSWAP": 3,
    "Toffoli": 3, "DoubleExcitation": 4, "DoubleExcitationMinus": 4, "DoubleExcitationPlus": 4,
    "MultiRZ": 3,
}


@pytest.mark.parametrize("name", sorted(GATE_ARITY))
def test_host_gate_compiler_matches_oracle(name):
    nw = GATE_ARITY[name]
    rng = np.random.default_rng(len(name))
    for _ in range(3):
        p = list(rng.uniform(-3, 3, N.PARAM_COUNT.get(name, 0)))
        for inv in (False, True):
            np.testing.assert_allclose(N.gate_matrix(name, list(range(nw)), p, inv),
                                       po.gate_matrix(name, nw, p, inv), atol=1e-14)


def test_host_gate_compiler_matrix_gate():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    np.testing.assert_allclose(N.gate_matrix("QubitUnitary", [0, 1, 2], [], False, A), A)
    np.testing.assert_allclose(N.gate_matrix("QubitUnitary", [0, 1, 2], [], True, A), A.conj().T)


def test_host_validation_errors():
    with pytest.raises(N.LghError):
        N.gate_matrix("CNOT", [0], [])          # wrong arity
    with pytest.raises(N.LghError):
        N.gate_matrix("CNOT", [1, 1], [])       # repeated wire
    with pytest.raises(ValueError):
        N.make_ops([("NotAGate", [0], [], False)])


REFERENCE_NAMES = [
    "LightningGPU_C64", "LightningGPU_C128", "NamedObsGPU_C64", "NamedObsGPU_C128", "HermitianObsGPU_C64",
    "HermitianObsGPU_C128", "TensorProdObsGPU_C64", "TensorProdObsGPU_C128", "HamiltonianGPU_C64",
    "HamiltonianGPU_C128", "SparseHamiltonianGPU_C64", "SparseHamiltonianGPU_C128", "OpsStructGPU_C64",
    "OpsStructGPU_C128", "AdjointJacobianGPU_C64", "AdjointJacobianGPU_C128", "DevPool", "DevTag",
    "PLException", "device_reset", "allToAllAccess", "is_gpu_supported", "get_gpu_arch",
]
SV_METHODS = [
    "setBasisState", "setStateVector", "Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT",
    "SWAP", "CSWAP", "Toffoli", "CY", "CZ", "PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "Rot", "CRX",
    "CRY", "CRZ", "CRot", "IsingXX", "IsingYY", "IsingZZ", "SingleExcitation", "SingleExcitationMinus",
    "SingleExcitationPlus", "DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ",
    "apply", "ExpectationValue", "Probability", "GenerateSamples", "DeviceToDevice", "DeviceToHost",
    "HostToDevice", "GetNumGPUs", "getCurrentGPU", "numQubits", "dataLength", "resetGPU",
]


def test_pybind_module_surface():
    m = pytest.importorskip("pennylane_lightning_gpu.lightning_gpu_qubit_ops")
    for name in REFERENCE_NAMES:
        assert hasattr(m, name), name
    for cls in (m.LightningGPU_C64, m.LightningGPU_C128):
        for meth in SV_METHODS:
            assert hasattr(cls, meth), meth
    for p in ("C64", "C128"):
        adj = getattr(m, f"AdjointJacobianGPU_{p}")()
        for meth in ("create_ops_list", "adjoint_jacobian", "adjoint_jacobian_batched"):
            assert hasattr(adj, meth)
    assert m.DevPool.getTotalDevices() >= 0


def test_pybind_observables_without_device():
    m = pytest.importorskip("pennylane_lightning_gpu.lightning_gpu_qubit_ops")
    a = m.NamedObsGPU_C128("PauliX", [3])
    assert repr(a) == "PauliX[3]"
    assert a.get_wires() == [3]
    assert a == m.NamedObsGPU_C128("PauliX", [3])
    assert not (a == m.NamedObsGPU_C128("PauliY", [3]))
    T = m.TensorProdObsGPU_C128([a, m.NamedObsGPU_C128("PauliZ", [1])])
    assert repr(T) == "PauliX[3] @ PauliZ[1]"
    assert T.get_wires() == [1, 3]
    with pytest.raises(m.PLException):
        m.TensorProdObsGPU_C128([a, m.NamedObsGPU_C128("PauliZ", [3])])
    H = m.HamiltonianGPU_C128(np.array([0.5, 2.0]), [a, T])
    assert "Hamiltonian" in repr(H)
    M = m.HermitianObsGPU_C64(np.eye(4, dtype=np.complex64).ravel(), [0, 2])
    assert M.get_wires() == [0, 2]
    S = m.SparseHamiltonianGPU_C128(np.array([1.0 + 0j]), np.array([0], np.int64), np.array([0, 1, 1], np.int64), [0])
    assert S.get_wires() == [0]


def test_state_creation_without_device_raises():
    m = pytest.importorskip("pennylane_lightning_gpu.lightning_gpu_qubit_ops")
    if N.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(m.PLException):
        m.LightningGPU_C128(3)
    with pytest.raises(N.LghError):
        N.State(3)


def test_batch_template_validation_without_device():
    """Template compilation is host work: structure checks and parameter
    counts need no GPU."""
    from pennylane_lightning_gpu.workloads import ring_template

    for n in (8, 12, 20):
        ops, pidx = ring_template(n)
        t = N.Template(n, ops, pidx, [("Z", 0)])
        assert (t.num_params, t.num_qubits) == (9 * n, n)
    ops = [("RX", [0], [0.0], False), ("CRY", [0, 1], [0.0], True), ("RX", [1], [0.0], False)]
    assert N.Template(2, ops, [0, 1, 0], [("Y", 1)]).num_params == 2
    bad = [
        (21, [("RX", [0], [0.1], False)], [0], [("Z", 0)], "1 to 20 qubits"),
        (4, [("SWAP", [0, 1], [], False)], [-1], [("Z", 0)], "not supported in a batched circuit"),
        (4, [("Hadamard", [0], [], False)], [0], [("Z", 0)], "cannot take a circuit parameter"),
        (4, [("RX", [5], [0.1], False)], [0], [("Z", 0)], "wire index out of range"),
        (4, [("RX", [0], [0.1], False)], [0], [("Z", 7)], "observable wire out of range"),
        (4, [("RX", [0], [0.1], False)], [0], [("Z", 1), ("X", 1)], "repeated observable wire"),
        (12, [("RX", [0], [0.1], False)], [0], [("X", w) for w in range(6)], "more than one pass"),
    ]
    for n, ops, pidx, obs, msg in bad:
        with pytest.raises(N.LghError, match=msg):
            N.Template(n, ops, pidx, obs)
    with pytest.raises(ValueError):
        N.batch_run([N.Template(2, ops, [0, 1, 0], [("Z", 0)])], [np.zeros(3)])
