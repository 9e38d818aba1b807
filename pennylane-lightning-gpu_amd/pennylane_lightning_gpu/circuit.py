"""PennyLane-free circuit description for the ``LightningGPU`` device harness.

PennyLane is not a dependency of this package.  The device harness
(``lightning_gpu.py``) reads operations, observables, measurements and tapes
through the attributes PennyLane's own classes expose (``name``, ``wires``,
``parameters``, ``num_params``, ``base`` of an adjoint, ``matrix()``,
``diagonalizing_gates()``, ``obs`` / ``coeffs`` / ``ops`` of tensors and
Hamiltonians, ``return_type`` of measurements, ``trainable_params`` and
``get_operation`` of tapes), so PennyLane objects work where PennyLane is
installed, and the small classes below stand in for them where it is not.

Conventions follow PennyLane (and the reference device,
pennylane_lightning_gpu/lightning_gpu.py): wire order is the order of the
wires given, the first wire is the most significant bit of a matrix index,
Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi).
"""
from __future__ import annotations

import math
from typing import Iterable, List, Sequence

import numpy as np

# --- gate and observable matrices (PennyLane definitions) --------------------

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)
_NAMED_MATRICES = {"Identity": _I2, "PauliX": _X, "PauliY": _Y, "PauliZ": _Z, "Hadamard": _H}


def _rz(phi):
    return np.diag([np.exp(-0.5j * phi), np.exp(0.5j * phi)])


def _ry(theta):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def embed(mat: np.ndarray, op_wires: Sequence, all_wires: Sequence) -> np.ndarray:
    """Matrix of ``mat`` (acting on ``op_wires``, first wire most significant)
    on the ordered wires ``all_wires``, identity on the others."""
    op_wires, all_wires = list(op_wires), list(all_wires)
    k = len(all_wires)
    rest = [w for w in all_wires if w not in op_wires]
    order = op_wires + rest
    full = np.kron(np.asarray(mat, dtype=np.complex128), np.eye(1 << len(rest), dtype=np.complex128))
    perm = [order.index(w) for w in all_wires]
    t = full.reshape([2] * (2 * k)).transpose(perm + [p + k for p in perm])
    return t.reshape(1 << k, 1 << k)


# --- operations ---------------------------------------------------------------

class Operation:
    """A gate: name, wires, parameters and, for gates without a built-in
    kernel (QubitUnitary and the like), its matrix."""

    def __init__(self, name: str, wires, params: Iterable = (), matrix=None):
        self.name = name
        self.wires = list(wires) if isinstance(wires, (list, tuple, range)) else [wires]
        self.parameters = list(params)
        self._matrix = None if matrix is None else np.asarray(matrix, dtype=np.complex128)

    @property
    def num_params(self) -> int:
        return len(self.parameters)

    def matrix(self) -> np.ndarray:
        if self._matrix is not None:
            return self._matrix
        if self.name in _NAMED_MATRICES:
            return _NAMED_MATRICES[self.name]
        if self.name == "RZ":
            return _rz(self.parameters[0])
        if self.name == "RY":
            return _ry(self.parameters[0])
        if self.name == "S":
            return np.diag([1, 1j]).astype(np.complex128)
        if self.name == "PhaseShift":
            return np.diag([1, np.exp(1j * self.parameters[0])])
        raise AttributeError(f"no matrix for {self.name}")

    def __repr__(self):
        return f"{self.name}({', '.join(map(str, self.parameters))}, wires={self.wires})"


class Adjoint(Operation):
    """The adjoint of ``base``; its matrix is the conjugate transpose."""

    def __init__(self, base: Operation):
        super().__init__("Adjoint(" + base.name + ")", base.wires, base.parameters)
        self.base = base

    def matrix(self) -> np.ndarray:
        return np.conj(self.base.matrix()).T


class Rot(Operation):
    """Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi)."""

    def __init__(self, phi, theta, omega, wires):
        super().__init__("Rot", wires, (phi, theta, omega))

    def expand(self):
        phi, theta, omega = self.parameters
        return _Expansion([Operation("RZ", self.wires, [phi]), Operation("RY", self.wires, [theta]),
                           Operation("RZ", self.wires, [omega])])


class _Expansion:
    def __init__(self, ops):
        self.operations = ops


class StatePrep(Operation):
    """Amplitudes of the listed wires (length 2^len(wires), normalised)."""

    def __init__(self, state, wires):
        super().__init__("StatePrep", wires, [np.asarray(state)])


class BasisState(Operation):
    """Computational basis state (0/1 per listed wire)."""

    def __init__(self, bits, wires):
        super().__init__("BasisState", wires, [np.asarray(bits)])


# --- observables --------------------------------------------------------------

class Observable:
    """Named observable (PauliX, PauliY, PauliZ, Hadamard, Identity)."""

    def __init__(self, name: str, wires):
        self.name = name
        self.wires = list(wires) if isinstance(wires, (list, tuple, range)) else [wires]
        self.parameters: List = []
        self._pauli_rep = None

    def matrix(self) -> np.ndarray:
        return _NAMED_MATRICES[self.name]

    def diagonalizing_gates(self) -> List[Operation]:
        w = self.wires
        if self.name == "PauliX":
            return [Operation("Hadamard", w)]
        if self.name == "PauliY":
            return [Operation("PauliZ", w), Operation("S", w), Operation("Hadamard", w)]
        if self.name == "Hadamard":
            return [Operation("RY", w, [-math.pi / 4])]
        if self.name in ("PauliZ", "Identity"):
            return []
        raise ValueError(f"no diagonalizing gates for {self.name}")

    def eigvals(self) -> np.ndarray:
        return np.array([1.0, 1.0]) if self.name == "Identity" else np.array([1.0, -1.0])

    def __matmul__(self, other):
        return Tensor(self, other)

    def __repr__(self):
        return f"{self.name}(wires={self.wires})"


def PauliX(w):  # noqa: N802 -- PennyLane's names
    return Observable("PauliX", w)


def PauliY(w):  # noqa: N802
    return Observable("PauliY", w)


def PauliZ(w):  # noqa: N802
    return Observable("PauliZ", w)


def Hadamard(w):  # noqa: N802
    return Observable("Hadamard", w)


def Identity(w):  # noqa: N802
    return Observable("Identity", w)


class Hermitian(Observable):
    """Hermitian matrix on the listed wires."""

    def __init__(self, matrix, wires):
        super().__init__("Hermitian", wires)
        self._mat = np.asarray(matrix, dtype=np.complex128)
        self.parameters = [self._mat]

    def matrix(self) -> np.ndarray:
        return self._mat


class Projector(Observable):
    """Projector onto a computational basis state of the listed wires."""

    def __init__(self, basis_state, wires):
        super().__init__("Projector", wires)
        self.basis_state = [int(b) for b in basis_state]
        self.parameters = [np.asarray(self.basis_state)]

    def matrix(self) -> np.ndarray:
        m = np.zeros((1 << len(self.wires), 1 << len(self.wires)), dtype=np.complex128)
        idx = int("".join(map(str, self.basis_state)), 2) if self.basis_state else 0
        m[idx, idx] = 1.0
        return m


class Tensor(Observable):
    """Product of observables on disjoint wires."""

    def __init__(self, *obs):
        flat = []
        for o in obs:
            flat.extend(o.obs if isinstance(o, Tensor) else [o])
        self.obs = flat
        super().__init__("Tensor", [w for o in flat for w in o.wires])
        self.name = [o.name for o in flat]

    @property
    def non_identity_obs(self):
        return [o for o in self.obs if o.name != "Identity"]

    def matrix(self) -> np.ndarray:
        m = np.ones((1, 1), dtype=np.complex128)
        for o in self.obs:
            m = np.kron(m, o.matrix())
        return m

    def diagonalizing_gates(self):
        return [g for o in self.obs for g in o.diagonalizing_gates()]

    def eigvals(self):
        e = np.ones(1)
        for o in self.obs:
            e = np.kron(e, o.eigvals())
        return e

    def __matmul__(self, other):
        return Tensor(self, other)


class Hamiltonian(Observable):
    """sum_k coeffs[k] * ops[k]; wires in order of first appearance."""

    def __init__(self, coeffs, ops):
        wires = []
        for o in ops:
            for w in o.wires:
                if w not in wires:
                    wires.append(w)
        super().__init__("Hamiltonian", wires)
        self.coeffs = np.asarray(coeffs, dtype=np.float64).reshape(-1)
        self.ops = list(ops)
        if len(self.coeffs) != len(self.ops):
            raise ValueError("number of coefficients and operators does not match")

    def matrix(self) -> np.ndarray:
        m = np.zeros((1 << len(self.wires), 1 << len(self.wires)), dtype=np.complex128)
        for c, o in zip(self.coeffs, self.ops):
            m += c * embed(o.matrix(), o.wires, self.wires)
        return m


class SparseHamiltonian(Observable):
    """Hamiltonian given as a scipy sparse matrix over the listed wires."""

    def __init__(self, H, wires):
        super().__init__("SparseHamiltonian", wires)
        self.H = H
        self.parameters = [H]

    def sparse_matrix(self):
        return self.H

    def matrix(self):
        return np.asarray(self.H.todense())


# --- measurements and tapes -------------------------------------------------------

Expectation, Variance, Probability, Sample, State = "expval", "var", "probs", "sample", "state"


class MeasurementProcess:
    def __init__(self, return_type: str, obs=None, wires=None):
        self.return_type = return_type
        self.obs = obs
        self.wires = list(wires) if wires is not None else (obs.wires if obs is not None else None)


def expval(obs):
    return MeasurementProcess(Expectation, obs)


def var(obs):
    return MeasurementProcess(Variance, obs)


def probs(wires=None):
    return MeasurementProcess(Probability, None, wires)


def sample(obs):
    return MeasurementProcess(Sample, obs)


def state():
    return MeasurementProcess(State)


class QuantumTape:
    """Operations and measurements; parameters are the operations' parameters
    in order (a StatePrep / BasisState array counts as one), and
    ``trainable_params`` indexes them (default: all)."""

    def __init__(self, operations, measurements, trainable_params=None):
        self.operations = list(operations)
        self.measurements = list(measurements)
        self._par_info = [(op, i) for op in self.operations for i in range(op.num_params)]
        self.trainable_params = (list(range(len(self._par_info))) if trainable_params is None
                                 else sorted(int(t) for t in trainable_params))

    @property
    def observables(self):
        return [m.obs for m in self.measurements]

    @property
    def num_params(self):
        return len(self.trainable_params)

    def get_operation(self, idx):
        """(operation, operation index, parameter index within it) of the
        idx-th trainable parameter."""
        op, p = self._par_info[self.trainable_params[idx]]
        return op, self.operations.index(op), p

    def copy(self):
        t = QuantumTape(self.operations, self.measurements, self.trainable_params)
        return t
