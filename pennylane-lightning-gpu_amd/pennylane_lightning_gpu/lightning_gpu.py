"""``LightningGPU``: the device harness above ``lightning_gpu_qubit_ops``.

Reproduces the routing of the reference device class
(pennylane_lightning_gpu/lightning_gpu.py:206-960) without importing
PennyLane.  Operations, observables, measurements and tapes are read
duck-typed (PennyLane objects, or the stand-ins of ``circuit.py``):

- ``apply``: a leading StatePrep / BasisState initialises the state
  (lightning_gpu.py:391-467); one anywhere else is an error; Identity is
  skipped; an adjoint runs its base with the inverse flag; a gate without a
  named method on the state-vector class runs by its matrix (:519-557).
- ``expval`` (:820-891): SparseHamiltonian through the resident-CSR SpMV;
  Hamiltonian as one dense matrix when it spans fewer than 14 wires, else as
  Pauli words; any other observable through its matrix; with shots, the mean
  of samples.  Hermitian and Projector use their matrix here (the reference
  defers them to PennyLane's generic eigenvalue path; the value is the same).
- ``var`` (:936-960): <M^dagger M> - <M>^2, both through matrices.
- ``probability`` (:893-924): wires must be in increasing order; the
  backend's reversed bit order is transposed back.
- ``adjoint_jacobian`` (:587-758) and ``vjp`` (:771-810): the checks on
  measurements and operations, serialization (_serialize.py), trainable
  parameter shifts for a leading state preparation, the optional observable
  batching, and the row reduction of split Hamiltonians.

Deliberate differences from the reference (each a reference quirk):
``starting_state`` is loaded into the device state (the reference reads it
into a variable it never uses); Pauli-word wires of a Hamiltonian go through
the wire map (the reference passes raw labels); sampling for an expectation
or variance rotates a copy of the state instead of the device state; the
Jacobian columns of trainable state-preparation parameters are zero (the
reference's reshape fails there).
"""
from __future__ import annotations

import warnings
from typing import Sequence, Union

import numpy as np

from . import _serialize as ser

_name_map = {"PauliX": "X", "PauliY": "Y", "PauliZ": "Z", "Identity": "I"}
_STATE_PREP = ("StatePrep", "QubitStateVector")
_PREP_ALL = _STATE_PREP + ("BasisState",)
_SHOTS_WARNING = ("Requested adjoint differentiation to be computed with finite shots. The derivative is always "
                  "exact when using the adjoint differentiation method.")


class DeviceError(Exception):
    """Invalid use of the device (mirrors pennylane.DeviceError)."""


class QuantumFunctionError(Exception):
    """Unsupported differentiation request (mirrors pennylane.QuantumFunctionError)."""


def _load_backend():
    from . import lightning_gpu_qubit_ops as backend  # the pybind11 module; no CPU fallback

    return backend


def _is_tensor(ob) -> bool:
    return hasattr(ob, "obs") and isinstance(getattr(ob, "name", None), list)


class LightningGPU:
    """State-vector device on one MI355X (lightning_gpu.py:206-378).

    Args:
        wires: number of wires, or the wire labels in device order.
        c_dtype: np.complex128 or np.complex64.
        shots: None for exact results, else the number of samples.
        batch_obs: False, True, or the number of observables per batch for
            the adjoint Jacobian (lightning_gpu.py:700-736).
        seed: seed of the sampler (the reference's sampler is unseeded).
    """

    short_name = "lightning.gpu"
    observables = {"PauliX", "PauliY", "PauliZ", "Hadamard", "SparseHamiltonian", "Hamiltonian", "Identity",
                   "Hermitian", "Projector"}

    def __init__(self, wires: Union[int, Sequence], *, c_dtype=np.complex128, shots=None,
                 batch_obs: Union[bool, int] = False, seed=None, _backend=None):
        if c_dtype is np.complex64:
            self.use_csingle, self.C_DTYPE, self.R_DTYPE = True, np.complex64, np.float32
        elif c_dtype is np.complex128:
            self.use_csingle, self.C_DTYPE, self.R_DTYPE = False, np.complex128, np.float64
        else:
            raise TypeError(f"Unsupported complex Type: {c_dtype}")
        self.wires = list(range(wires)) if isinstance(wires, int) else list(wires)
        if len(set(self.wires)) != len(self.wires):
            raise ValueError("wire labels must be unique")
        self.wire_map = {w: i for i, w in enumerate(self.wires)}
        self.num_wires = len(self.wires)
        self.shots = shots
        self._batch_obs = batch_obs
        self._backend = _backend if _backend is not None else _load_backend()
        self._gpu_state = getattr(self._backend, "LightningGPU_C" + ("64" if self.use_csingle else "128"))(
            self.num_wires)
        if seed is not None:
            self._gpu_state.setSeed(int(seed))
        self._samples = None
        self._create_basis_state_GPU(0)

    # --- wires and state -----------------------------------------------------------
    def map_wires(self, wires) -> list:
        try:
            return [self.wire_map[w] for w in wires]
        except KeyError as e:
            raise DeviceError(f"wire {e.args[0]} is not on the device") from None

    def reset(self):
        self._samples = None
        self._gpu_state.resetGPU(False)

    @property
    def state(self) -> np.ndarray:
        out = np.zeros(1 << self.num_wires, dtype=self.C_DTYPE)
        self.syncD2H(out)
        return out

    def syncD2H(self, state_vector, use_async=False):
        self._gpu_state.DeviceToHost(state_vector.ravel(order="C"), use_async)

    def syncH2D(self, state_vector, use_async=False):
        self._gpu_state.HostToDevice(np.ascontiguousarray(state_vector, dtype=self.C_DTYPE).ravel(order="C"),
                                     use_async)

    def _create_basis_state_GPU(self, index, use_async=False):
        self._gpu_state.setBasisState(int(index), use_async)

    def _apply_state_vector_GPU(self, state, device_wires, use_async=False):
        """lightning_gpu.py:391-439: the full state, or amplitudes scattered
        onto the basis states of a subset of wires (others |0>)."""
        device_wires = self.map_wires(device_wires)
        dim = 1 << len(device_wires)
        state = np.asarray(state, dtype=self.C_DTYPE).reshape(-1)
        if state.shape != (dim,):
            raise ValueError("State vector must have shape (2**wires,) or (batch_size, 2**wires).")
        if not np.allclose(np.linalg.norm(state), 1.0, atol=1e-6 if self.use_csingle else 1e-10):
            raise ValueError("Sum of amplitudes-squared does not equal one.")
        if len(device_wires) == self.num_wires and sorted(device_wires) == device_wires:
            self.syncH2D(state)
            return
        basis = (np.arange(dim)[:, None] >> np.arange(len(device_wires) - 1, -1, -1)[None, :]) & 1
        index = np.zeros(dim, dtype=np.int64)
        for j, w in enumerate(device_wires):
            index |= basis[:, j].astype(np.int64) << (self.num_wires - 1 - w)
        self._gpu_state.resetGPU(False)
        idx_t = np.int32 if self.use_csingle else np.int64
        self._gpu_state.setStateVector(index.astype(idx_t), state, use_async)

    def _apply_basis_state_GPU(self, state, wires):
        """lightning_gpu.py:441-467."""
        device_wires = self.map_wires(wires)
        state = list(np.asarray(state).reshape(-1).tolist())
        if not set(state).issubset({0, 1}):
            raise ValueError("BasisState parameter must consist of 0 or 1 integers.")
        if len(state) != len(device_wires):
            raise ValueError("BasisState parameter and wires must be of equal length.")
        num = sum(int(b) << (self.num_wires - 1 - w) for b, w in zip(state, device_wires))
        self._create_basis_state_GPU(num)

    # --- operations ----------------------------------------------------------------
    def apply_cq(self, operations, **kwargs):
        """lightning_gpu.py:519-555."""
        for o in operations:
            if o.name == "Identity":
                continue
            invert = hasattr(o, "base")
            name = o.base.name if invert else o.name
            method = getattr(self._gpu_state, name, None)
            wires = self.map_wires(o.wires)
            if method is None:
                mat = np.asarray(o.matrix(), dtype=self.C_DTYPE)  # already holds the adjoint
                if mat.size == 0:
                    raise DeviceError("Unsupported operation")
                self._gpu_state.apply(name, wires, False, [], mat.ravel(order="C"))
            else:
                method(wires, invert, [float(p) for p in np.ravel(o.parameters)] if o.parameters else [])

    def apply(self, operations, **kwargs):
        """lightning_gpu.py:557-575."""
        operations = list(operations)
        if operations:
            if operations[0].name in _STATE_PREP:
                self._apply_state_vector_GPU(np.asarray(operations[0].parameters[0]).copy(), operations[0].wires)
                operations = operations[1:]
            elif operations[0].name == "BasisState":
                self._apply_basis_state_GPU(operations[0].parameters[0], operations[0].wires)
                operations = operations[1:]
        for op in operations:
            if op.name in _STATE_PREP or op.name == "BasisState":
                raise DeviceError(f"Operation {op.name} cannot be used after other Operations have already been "
                                  f"applied on a {self.short_name} device.")
        self.apply_cq(operations)

    # --- measurements --------------------------------------------------------------
    def expval(self, observable, shot_range=None, bin_size=None):
        """lightning_gpu.py:820-891."""
        if self.shots is not None:
            return float(np.squeeze(np.mean(self.sample(observable), axis=0)))
        if observable.name == "SparseHamiltonian":
            csr = observable.sparse_matrix().tocsr()
            idx_t = np.int32 if self.use_csingle else np.int64
            return self._gpu_state.ExpectationValue(csr.indptr.astype(idx_t), csr.indices.astype(idx_t),
                                                    np.asarray(csr.data, dtype=self.C_DTYPE))
        if observable.name == "Hamiltonian":
            device_wires = self.map_wires(observable.wires)
            if len(device_wires) < 14:
                return self._gpu_state.ExpectationValue(device_wires,
                                                        np.asarray(observable.matrix(), self.C_DTYPE).ravel())
            pauli_words, word_wires = [], []
            for word in observable.ops:
                names = word.name if isinstance(word.name, list) else [word.name]
                compressed = []
                for nm in names:
                    if nm not in _name_map:
                        raise ValueError("Pauli word only for Hamiltionian expval.")
                    compressed.append(_name_map[nm])
                word_wires.append(self.map_wires(word.wires))
                pauli_words.append("".join(compressed))
            return self._gpu_state.ExpectationValue(pauli_words, word_wires,
                                                    np.asarray(observable.coeffs, dtype=self.C_DTYPE))
        mat = np.asarray(observable.matrix(), dtype=self.C_DTYPE).ravel(order="C")
        wires = self.map_wires(observable.wires)
        if isinstance(observable.name, list):
            return self._gpu_state.ExpectationValue(list(observable.name), wires, [], mat)
        return self._gpu_state.ExpectationValue(observable.name, wires, [], mat)

    def var(self, observable, shot_range=None, bin_size=None):
        """lightning_gpu.py:936-960."""
        if self.shots is not None:
            return float(np.squeeze(np.var(self.sample(observable), axis=0)))
        m = np.asarray(observable.matrix(), dtype=np.complex128)
        sqr = np.conj(m).T @ m
        names = observable.name if isinstance(observable.name, list) else [observable.name]
        wires = self.map_wires(observable.wires)
        mean = self._gpu_state.ExpectationValue([n + "_var" for n in names], wires, [],
                                                m.astype(self.C_DTYPE).ravel())
        squared_mean = self._gpu_state.ExpectationValue([n + "_sqr" for n in names], wires, [],
                                                        sqr.astype(self.C_DTYPE).ravel())
        return squared_mean - mean ** 2

    def probability(self, wires=None, shot_range=None, bin_size=None):
        """lightning_gpu.py:893-924."""
        wires = list(self.wires) if wires is None else list(wires)
        device_wires = self.map_wires(wires)
        if self.shots is not None:
            return self.estimate_probability(wires)
        if len(device_wires) > 1 and not all(a <= b for a, b in zip(device_wires[:-1], device_wires[1:])):
            raise RuntimeError("Lightning does not currently support out-of-order indices for probabilities")
        local_prob = np.asarray(self._gpu_state.Probability(device_wires))
        if local_prob.size == 0:
            return local_prob
        k = local_prob.size.bit_length() - 1
        return local_prob.reshape([2] * k).transpose().reshape(-1)

    def generate_samples(self):
        """(shots, num_wires) array of bits (lightning_gpu.py:926-932)."""
        return np.asarray(self._gpu_state.GenerateSamples(self.num_wires, int(self.shots))).astype(int)

    def sample(self, observable, shot_range=None, bin_size=None):
        """Eigenvalue samples of a Pauli-product / Hadamard / Identity
        observable: its diagonalizing gates act on a copy of the state."""
        if self.shots is None:
            raise DeviceError("sampling needs a finite number of shots")
        gates = observable.diagonalizing_gates()
        if gates:
            saved = self._gpu_state
            self._gpu_state = type(saved)(saved)
            try:
                self.apply_cq(gates)
                samples = self.generate_samples()
            finally:
                self._gpu_state = saved
        else:
            samples = self.generate_samples()
        self._samples = samples
        obs_list = observable.obs if _is_tensor(observable) else [observable]
        out = np.ones(samples.shape[0])
        for o in obs_list:
            if o.name == "Identity":
                continue
            for w in self.map_wires(o.wires):
                out = out * (1 - 2 * samples[:, w])
        return out

    def estimate_probability(self, wires):
        samples = self.generate_samples()[:, self.map_wires(wires)]
        idx = samples @ (1 << np.arange(len(wires) - 1, -1, -1))
        return np.bincount(idx, minlength=1 << len(wires)) / samples.shape[0]

    def execute(self, tape):
        """Reset, apply the tape's operations, return its measurements."""
        self.reset()
        self.apply(tape.operations)
        out = []
        for m in tape.measurements:
            rt = m.return_type
            if rt == "expval":
                out.append(self.expval(m.obs))
            elif rt == "var":
                out.append(self.var(m.obs))
            elif rt == "probs":
                out.append(self.probability(m.wires))
            elif rt == "sample":
                out.append(self.sample(m.obs))
            elif rt == "state":
                out.append(self.state)
            else:
                raise DeviceError(f"unsupported measurement {rt}")
        return out[0] if len(out) == 1 else out

    # --- adjoint differentiation ----------------------------------------------------
    @staticmethod
    def _check_adjdiff_supported_measurements(measurements):
        """lightning_gpu.py:577-622."""
        if len(measurements) == 0:
            return None
        if len(measurements) == 1 and measurements[0].return_type == "state":
            raise QuantumFunctionError("Not supported")
        if not all(m.return_type == "expval" for m in measurements):
            raise QuantumFunctionError("Adjoint differentiation method does not support expectation return type "
                                       "mixed with other return types")
        for m in measurements:
            obs = m.obs.non_identity_obs if _is_tensor(m.obs) else [m.obs]
            if any(o.name == "Projector" for o in obs):
                raise QuantumFunctionError("Adjoint differentiation method does not support the Projector observable")
            if any(o.name == "Hermitian" for o in obs):
                raise QuantumFunctionError("LightningGPU adjoint differentiation method does not currently support "
                                           "the Hermitian observable")
        return "expval"

    @staticmethod
    def _check_adjdiff_supported_operations(operations):
        """lightning_gpu.py:624-637."""
        for op in operations:
            if op.num_params > 1 and op.name != "Rot":
                raise QuantumFunctionError(f"The {op.name} operation is not supported using "
                                           'the "adjoint" differentiation method')

    def _adj(self):
        return getattr(self._backend, "AdjointJacobianGPU_C" + ("64" if self.use_csingle else "128"))()

    def _serializer(self) -> ser.TapeSerializer:
        return ser.TapeSerializer(self._backend, self.wire_map, self.use_csingle)

    @staticmethod
    def _parameter_columns(tape, shift: int):
        """Which tape parameters the backend differentiates: for every
        trainable parameter that is not a state preparation's, its index
        among the parameters the backend sees (the serialized operations
        leave the state preparation's parameter out, ``shift``) and its
        column in the device's Jacobian (lightning_gpu.py:675-693)."""
        backend_index, column = [], []
        for col, par in enumerate(sorted(tape.trainable_params)):
            if tape.get_operation(col)[0].name in _PREP_ALL:
                continue
            backend_index.append(par - shift)
            column.append(col)
        return backend_index, column

    def _observable_batches(self, count: int):
        """Slices of the serialized observables run per backend call: all of
        them at once, or groups of ``batch_obs`` (lightning_gpu.py:700-721)."""
        if not self._batch_obs or count == 0:
            return None
        size = count if isinstance(self._batch_obs, bool) else max(1, int(self._batch_obs))
        return [slice(lo, min(lo + size, count)) for lo in range(0, count, size)]

    def adjoint_jacobian(self, tape, starting_state=None, use_device_state=False, **kwargs):
        """The Jacobian of the tape's expectation values by the adjoint
        method (lightning_gpu.py:639-758): one row per tape observable, one
        column per trainable parameter; a split sum's rows are added back."""
        if self.shots is not None:
            warnings.warn(_SHOTS_WARNING, UserWarning)
        self._check_adjdiff_supported_measurements(tape.measurements)
        if not tape.trainable_params:
            return np.array(0)
        self._check_adjdiff_supported_operations(tape.operations)
        if starting_state is not None:
            self.syncH2D(np.ravel(starting_state, order="C"))
        elif not use_device_state:
            self.reset()
            self.apply(tape.operations)

        serializer = self._serializer()
        rows, offsets = serializer.observables(tape)
        op_lists, has_prep = serializer.operations(tape)
        adj = self._adj()
        ops = adj.create_ops_list(*op_lists)
        params, columns = self._parameter_columns(tape, 1 if has_prep else 0)

        batches = self._observable_batches(len(rows))
        if batches is None:
            raw = adj.adjoint_jacobian(self._gpu_state, rows, ops, params)
        else:
            raw = [r for b in batches for r in adj.adjoint_jacobian_batched(self._gpu_state, rows[b], ops, params)]
        per_row = np.asarray(raw, dtype=np.float64).reshape(len(rows), len(params))
        # rows of one tape observable (a split sum has several) add up
        per_ob = np.stack([per_row[lo:hi].sum(axis=0) for lo, hi in zip(offsets[:-1], offsets[1:])]) \
            if len(offsets) > 1 else np.zeros((0, len(params)))
        full = np.zeros((len(offsets) - 1, len(tape.trainable_params)))
        full[:, columns] = per_ob
        return self._adjoint_jacobian_processing(full)

    @staticmethod
    def _adjoint_jacobian_processing(jac):
        """The return layout of the reference (lightning_gpu.py:755-768): the
        squeezed Jacobian as a 0-d array, or as (nested) tuples of 0-d
        arrays."""
        jac = np.squeeze(jac)

        def nest(a):
            return np.array(a) if a.ndim == 0 else tuple(nest(x) for x in a)

        return nest(jac)

    def vjp(self, measurements, dy, starting_state=None, use_device_state=False):
        """Vector-Jacobian product dy . J (lightning_gpu.py:771-810), returned
        as a function of the tape: the Jacobian of the single expectation
        <sum_j dy_j O_j>, one adjoint sweep instead of one per observable."""
        from .circuit import Hamiltonian, MeasurementProcess

        if self.shots is not None:
            warnings.warn(_SHOTS_WARNING, UserWarning)
        kind = self._check_adjdiff_supported_measurements(measurements)
        dy = np.asarray(dy)
        if kind is None or not np.any(dy):
            return lambda tape: np.zeros(len(tape.trainable_params))
        if dy.shape[0] != len(measurements):
            raise ValueError(f"dy has {dy.shape[0]} entries for {len(measurements)} measured observables; the "
                             "vjp needs one per observable")
        if np.iscomplexobj(dy):
            raise ValueError("the vjp of expectation values takes a real dy")
        weighted = Hamiltonian(dy, [m.obs for m in measurements])

        def processing_fn(tape):
            if not tape.trainable_params:
                return np.array([], dtype=self.C_DTYPE)
            reweighted = tape.copy()
            reweighted.measurements = [MeasurementProcess("expval", weighted)]
            return self.adjoint_jacobian(reweighted, starting_state, use_device_state)

        return processing_fn
