"""Serialization of tapes into the objects of ``lightning_gpu_qubit_ops``.

Same functions, argument meaning and results as the reference's
pennylane_lightning_gpu/_serialize.py:142-336 (observables and operations for
the adjoint Jacobian), with the backend module passed in instead of imported:
``backend`` is the pybind11 module (or a stand-in with the same class names).

- Observables: named, tensor, Hamiltonian (split into one single-term
  Hamiltonian per term unless ``use_splitting`` is False; the device sums the
  rows back), SparseHamiltonian (CSR), Pauli sentences.  Hermitian
  observables are rejected for the adjoint method, as in the reference.
- Operations: StatePrep / BasisState are skipped (and reported); Rot expands
  into RZ, RY, RZ; an adjoint becomes its base with the inverse flag; a gate
  without a named method on the state-vector class goes by its matrix (which
  already includes the adjoint, so its inverse flag is cleared).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

pauli_name_map = {"I": "Identity", "X": "PauliX", "Y": "PauliY", "Z": "PauliZ"}


def _bits(use_csingle: bool) -> str:
    return "64" if use_csingle else "128"


def _named_ob_dtype(backend, use_csingle):
    return getattr(backend, "NamedObsGPU_C" + _bits(use_csingle))


def _tensor_ob_dtype(backend, use_csingle):
    return getattr(backend, "TensorProdObsGPU_C" + _bits(use_csingle))


def _hamiltonian_ob_dtype(backend, use_csingle):
    return getattr(backend, "HamiltonianGPU_C" + _bits(use_csingle)), (np.float32 if use_csingle else np.float64)


def _sparsehamiltonian_ob_dtype(backend, use_csingle):
    return (getattr(backend, "SparseHamiltonianGPU_C" + _bits(use_csingle)),
            np.complex64 if use_csingle else np.complex128, np.int32 if use_csingle else np.int64)


def _sv_py_dtype(backend, use_csingle):
    return getattr(backend, "LightningGPU_C" + _bits(use_csingle))


def _is_tensor(ob) -> bool:
    return hasattr(ob, "obs") and isinstance(getattr(ob, "name", None), list)


def _serialize_named_ob(o, wires_map: dict, backend, use_csingle: bool):
    """_serialize.py:142-149."""
    wires = [wires_map[w] for w in o.wires]
    if o.name == "Identity":
        wires = wires[:1]
    return _named_ob_dtype(backend, use_csingle)(o.name, wires)


def _serialize_tensor_ob(ob, wires_map: dict, backend, use_csingle: bool):
    """_serialize.py:152-156."""
    return _tensor_ob_dtype(backend, use_csingle)(
        [_serialize_ob(o, wires_map, backend, use_csingle) for o in ob.obs])


def _serialize_hamiltonian(ob, wires_map: dict, backend, use_csingle: bool, split_terms: bool = True):
    """_serialize.py:159-168."""
    hamiltonian_obs, rtype = _hamiltonian_ob_dtype(backend, use_csingle)
    coeffs = np.array(ob.coeffs).astype(rtype)
    terms = [_serialize_ob(t, wires_map, backend, use_csingle) for t in ob.ops]
    if split_terms:
        return [hamiltonian_obs(np.array([c], dtype=rtype), [t]) for (c, t) in zip(coeffs, terms)]
    return hamiltonian_obs(coeffs, terms)


def _serialize_sparsehamiltonian(ob, wires_map: dict, backend, use_csingle: bool):
    """_serialize.py:171-193 (single process)."""
    sparse_obs, ctype, rtype = _sparsehamiltonian_ob_dtype(backend, use_csingle)
    spm = ob.sparse_matrix().tocsr()
    data = np.array(spm.data).astype(ctype)
    indices = np.array(spm.indices).astype(rtype)
    offsets = np.array(spm.indptr).astype(rtype)
    return sparse_obs(data, indices, offsets, [wires_map[w] for w in ob.wires])


def _serialize_pauli_word(ob, wires_map: dict, backend, use_csingle: bool):
    """_serialize.py:204-215: a {wire: 'X'|'Y'|'Z'|'I'} word."""
    named_obs = _named_ob_dtype(backend, use_csingle)
    items = list(ob.items())
    if len(items) == 1:
        wire, pauli = items[0]
        return named_obs(pauli_name_map[pauli], [wires_map[wire]])
    return _tensor_ob_dtype(backend, use_csingle)(
        [named_obs(pauli_name_map[pauli], [wires_map[wire]]) for wire, pauli in items])


def _serialize_pauli_sentence(ob, wires_map: dict, backend, use_csingle: bool, split_terms: bool = True):
    """_serialize.py:218-229: a {word: coefficient} sentence."""
    hamiltonian_obs, rtype = _hamiltonian_ob_dtype(backend, use_csingle)
    pwords, coeffs = zip(*ob.items())
    terms = [_serialize_pauli_word(pw, wires_map, backend, use_csingle) for pw in pwords]
    coeffs = np.array(coeffs).astype(rtype)
    if split_terms:
        return [hamiltonian_obs(np.array([c], dtype=rtype), [t]) for (c, t) in zip(coeffs, terms)]
    return hamiltonian_obs(coeffs, terms)


def _serialize_ob(ob, wires_map, backend, use_csingle, use_splitting: bool = True):
    """_serialize.py:232-251."""
    if _is_tensor(ob):
        return _serialize_tensor_ob(ob, wires_map, backend, use_csingle)
    if ob.name == "Hamiltonian":
        return _serialize_hamiltonian(ob, wires_map, backend, use_csingle, use_splitting)
    if ob.name == "SparseHamiltonian":
        return _serialize_sparsehamiltonian(ob, wires_map, backend, use_csingle)
    if ob.name in ("PauliX", "PauliY", "PauliZ", "Identity", "Hadamard"):
        return _serialize_named_ob(ob, wires_map, backend, use_csingle)
    if getattr(ob, "_pauli_rep", None) is not None:
        return _serialize_pauli_sentence(ob._pauli_rep, wires_map, backend, use_csingle, use_splitting)
    if ob.name == "Hermitian":
        raise TypeError("Hermitian observables are not currently supported for adjoint differentiation. "
                        "Please use Pauli-words only.")
    raise TypeError(f"Unknown observable found: {ob}. Please use Pauli-words only.")


def _serialize_observables(tape, wires_map: dict, backend, use_csingle: bool = False) -> Tuple[List, List[int]]:
    """_serialize.py:254-279: serialized observables and, per tape
    observable, the offset of its rows (a split Hamiltonian takes one row per
    term)."""
    output = []
    offsets = [0]
    for ob in tape.observables:
        ser_ob = _serialize_ob(ob, wires_map, backend, use_csingle)
        if isinstance(ser_ob, list):
            output.extend(ser_ob)
            offsets.append(offsets[-1] + len(ser_ob))
        else:
            output.append(ser_ob)
            offsets.append(offsets[-1] + 1)
    return output, offsets


def _serialize_ops(tape, wires_map: dict, backend, use_csingle: bool = False):
    """_serialize.py:282-336: ((names, params, wires, inverses, matrices),
    uses_stateprep)."""
    names, params, wires, inverses, mats = [], [], [], [], []
    uses_stateprep = False
    sv_py = _sv_py_dtype(backend, use_csingle)
    rtype = np.float32 if use_csingle else np.float64
    ctype = np.complex64 if use_csingle else np.complex128
    for o in tape.operations:
        if o.name in ("StatePrep", "QubitStateVector", "BasisState"):
            uses_stateprep = True
            continue
        op_list = o.expand().operations if o.name == "Rot" else [o]
        for single_op in op_list:
            is_inverse = hasattr(single_op, "base")
            name = single_op.base.name if is_inverse else single_op.name
            names.append(name)
            if getattr(sv_py, name, None) is None:
                params.append(np.zeros(0, dtype=rtype))
                mats.append(np.asarray(single_op.matrix(), dtype=ctype).ravel(order="C"))
                is_inverse = False  # the matrix already holds the adjoint
            else:
                params.append(np.asarray(single_op.parameters, dtype=rtype))
                mats.append(np.zeros(0, dtype=ctype))
            wires.append([wires_map[w] for w in single_op.wires])
            inverses.append(is_inverse)
    return (names, params, wires, inverses, mats), uses_stateprep
