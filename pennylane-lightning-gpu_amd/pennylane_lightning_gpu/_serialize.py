"""Tape serialization for the adjoint Jacobian of ``LightningGPU``.

``TapeSerializer`` holds the backend classes of one precision and the
device's wire map, and turns a tape into the arguments of
``AdjointJacobianGPU_*.create_ops_list`` and ``adjoint_jacobian``:

- ``observables(tape)``: backend observable objects for all tape
  observables, flattened, with the offset of each tape observable's first
  row.  A sum (Hamiltonian, or a Pauli sentence) gives one single-term
  Hamiltonian per term when sums are split, so the adjoint sweep computes
  one Jacobian row per term and the device adds the rows back.
- ``operations(tape)``: the five OpsData lists -- names, parameter arrays,
  wires, inverse flags, matrices -- and whether the tape holds a state
  preparation.

The objects built and their arguments are those of the reference's
serializer (pennylane_lightning_gpu/_serialize.py:142-336), which is the
wire format of the pybind11 boundary:

- named observables take mapped wires, Identity only its first wire;
- a tensor product is built from its factors, a Pauli word {wire: letter}
  from one named observable per letter (the named observable itself for a
  single letter);
- a sparse Hamiltonian is passed as CSR (data, indices, indptr) in the
  precision's complex and index types, with its wires mapped;
- Hermitian observables are refused for the adjoint method;
- StatePrep / QubitStateVector / BasisState are left out and reported;
  Rot becomes RZ(phi), RY(theta), RZ(omega); an adjoint runs its base with
  the inverse flag; an operation without a named method on the state-vector
  class is passed by its matrix, which already holds any adjoint.

Objects are read duck-typed, so PennyLane objects and the stand-ins of
``circuit.py`` both fit.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

import numpy as np

_LETTER = {"I": "Identity", "X": "PauliX", "Y": "PauliY", "Z": "PauliZ"}
_NAMED = frozenset(("PauliX", "PauliY", "PauliZ", "Identity", "Hadamard"))
_PREP = frozenset(("StatePrep", "QubitStateVector", "BasisState"))
# gates the adjoint sweep takes as products of one-parameter gates:
# (gate, index of the parameter it takes), in application order
_PRODUCTS = {"Rot": (("RZ", 0), ("RY", 1), ("RZ", 2))}


@dataclass(frozen=True)
class _Precision:
    suffix: str
    real: type
    cplx: type
    index: type


_PRECISION = {
    False: _Precision("C128", np.float64, np.complex128, np.int64),
    True: _Precision("C64", np.float32, np.complex64, np.int32),
}


def _is_product(ob) -> bool:
    return isinstance(getattr(ob, "name", None), list) and hasattr(ob, "obs")


class TapeSerializer:
    """Serializes tapes for one backend precision (``single``: complex64).

    ``split_sums``: a sum observable gives one row per term (the adjoint
    method's default) instead of one row for the whole sum.
    """

    def __init__(self, backend, wire_map: Dict, single: bool = False, split_sums: bool = True):
        self.precision = _PRECISION[bool(single)]
        self.wire_map = wire_map
        self.split_sums = split_sums
        suffix = self.precision.suffix

        def cls(stem):
            return getattr(backend, f"{stem}_{suffix}", None)

        self._named = cls("NamedObsGPU")
        self._product = cls("TensorProdObsGPU")
        self._sum = cls("HamiltonianGPU")
        self._sparse = cls("SparseHamiltonianGPU")
        self.state_class = cls("LightningGPU")

    def _map(self, wires) -> List[int]:
        return [self.wire_map[w] for w in wires]

    # ---------------------------------------------------------------- observables
    def _classify(self, ob) -> str:
        if _is_product(ob):
            return "product"
        name = getattr(ob, "name", None)
        if name in ("Hamiltonian", "SparseHamiltonian"):
            return name
        if name in _NAMED:
            return "named"
        if getattr(ob, "_pauli_rep", None) is not None:
            return "sentence"
        if name == "Hermitian":
            raise TypeError("Hermitian observables cannot be differentiated with the adjoint method on this "
                            "device; express the observable through Pauli words.")
        raise TypeError(f"Observable {ob} is not supported by the adjoint method; express it through Pauli words.")

    def _named_of(self, name: str, wires: Sequence) -> object:
        mapped = self._map(wires)
        return self._named(name, mapped[:1] if name == "Identity" else mapped)

    def _word(self, word: Dict) -> object:
        """A Pauli word {wire: letter}."""
        factors = [self._named_of(_LETTER[letter], [wire]) for wire, letter in word.items()]
        return factors[0] if len(factors) == 1 else self._product(factors)

    def _sum_terms(self, ob, kind: str) -> Tuple[np.ndarray, List]:
        """(coefficients, term objects) of a sum observable."""
        if kind == "Hamiltonian":
            coeffs, terms = list(ob.coeffs), [self.single(t) for t in ob.ops]
        else:
            items = list(ob._pauli_rep.items())
            coeffs, terms = [c for _, c in items], [self._word(w) for w, _ in items]
        return np.asarray(coeffs).astype(self.precision.real), terms

    def single(self, ob) -> object:
        """One backend object for ``ob`` (a sum stays whole)."""
        kind = self._classify(ob)
        if kind == "product":
            return self._product([self.single(o) for o in ob.obs])
        if kind == "named":
            return self._named_of(ob.name, ob.wires)
        if kind == "SparseHamiltonian":
            csr = ob.sparse_matrix().tocsr()
            p = self.precision
            return self._sparse(np.asarray(csr.data).astype(p.cplx), np.asarray(csr.indices).astype(p.index),
                                np.asarray(csr.indptr).astype(p.index), self._map(ob.wires))
        coeffs, terms = self._sum_terms(ob, kind)
        return self._sum(coeffs, terms)

    def rows(self, ob) -> List:
        """The backend objects standing for ``ob``, one per Jacobian row."""
        kind = self._classify(ob)
        if self.split_sums and kind in ("Hamiltonian", "sentence"):
            coeffs, terms = self._sum_terms(ob, kind)
            return [self._sum(coeffs[i:i + 1], [t]) for i, t in enumerate(terms)]
        return [self.single(ob)]

    def observables(self, tape) -> Tuple[List, List[int]]:
        """All rows of the tape's observables and the row offsets: the rows
        of observable k are rows[offsets[k]:offsets[k + 1]]."""
        per_ob = [self.rows(ob) for ob in tape.observables]
        offsets = [0]
        for r in per_ob:
            offsets.append(offsets[-1] + len(r))
        return [x for r in per_ob for x in r], offsets

    # ---------------------------------------------------------------- operations
    def _records(self, op) -> List[Tuple]:
        """(name, parameters, wires, inverse, matrix or None) per backend gate."""
        wires = self._map(op.wires)
        if op.name in _PRODUCTS:
            par = list(op.parameters)
            return [(g, [par[k]], wires, False, None) for g, k in _PRODUCTS[op.name]]
        base = getattr(op, "base", None)
        name = base.name if base is not None else op.name
        if getattr(self.state_class, name, None) is not None:
            return [(name, list(op.parameters), wires, base is not None, None)]
        return [(name, [], wires, False, np.asarray(op.matrix()))]

    def operations(self, tape):
        """((names, params, wires, inverses, matrices), holds_state_preparation)."""
        p = self.precision
        records, prep = [], False
        for op in tape.operations:
            if op.name in _PREP:
                prep = True
            else:
                records.extend(self._records(op))
        lists = ([], [], [], [], [])
        for name, par, wires, inverse, mat in records:
            lists[0].append(name)
            lists[1].append(np.asarray(par, dtype=p.real))
            lists[2].append(wires)
            lists[3].append(inverse)
            lists[4].append(np.zeros(0, dtype=p.cplx) if mat is None else np.asarray(mat, dtype=p.cplx).ravel(order="C"))
        return lists, prep
