"""ctypes front end of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path
never imports this module.

Operations are tuples ``(name, wires, params, inverse[, matrix])`` with
PennyLane conventions (wire 0 = most significant amplitude-index bit, matrices
row-major with the first wire as the most significant matrix-index bit).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_c128p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")
_c64p = np.ctypeslib.ndpointer(np.complex64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["sh", os.path.join(_HERE, "build.sh")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_gate_matrix.argtypes = [C.c_char_p, C.c_int, _f64p, C.c_int, _c128p]
        L.orc_generator.argtypes = [C.c_char_p, C.c_int, _c128p, C.POINTER(C.c_double)]
        L.orc_apply_ops.argtypes = [C.c_int, _c128p, C.c_int, C.POINTER(C.c_char_p), _i32p,
                                    _i32p, _i32p, _f64p, C.POINTER(C.c_void_p)]
        L.orc_apply_generator.argtypes = [C.c_int, _c128p, C.c_char_p, _i32p, C.c_int]
        L.orc_expval_matrix.argtypes = [C.c_int, _c128p, _i32p, C.c_int, _c128p,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_expval_pauli_words.argtypes = [C.c_int, _c128p, C.c_int, C.c_char_p, _i32p, _i64p,
                                             _c128p, _f64p, C.POINTER(C.c_double)]
        L.orc_spmv.argtypes = [C.c_int64, _i64p, _i64p, _c128p, _c128p, _c128p]
        L.orc_expval_csr.argtypes = [C.c_int, _c128p, _i64p, _i64p, _c128p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_probs.argtypes = [C.c_int, _c128p, _i32p, C.c_int, _f64p]
        L.orc_philox4x32_10.argtypes = [C.c_uint64, C.c_uint64, np.ctypeslib.ndpointer(np.uint32)]
        L.orc_philox4x32_10.restype = None
        L.orc_sample.argtypes = [C.c_int, _c128p, C.c_uint64, C.c_uint64, _u64p]
        L.orc_sample_c64.argtypes = [C.c_int, _c64p, C.c_uint64, C.c_uint64, _u64p]
        L.orc_adjoint_jacobian.argtypes = [C.c_int, _c128p, C.c_int, C.POINTER(C.c_void_p), C.c_int,
                                           C.POINTER(C.c_char_p), _i32p, _i32p, _i32p, _f64p, _i32p,
                                           C.POINTER(C.c_void_p), _i64p, C.c_int, _f64p]
        _lib = L
    return _lib


def num_threads() -> int:
    return lib().orc_num_threads()


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def zero_state(n: int) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def gate_matrix(name: str, nw: int, params=(), inverse=False) -> np.ndarray:
    p = np.zeros(3)
    p[: len(params)] = params
    out = np.zeros((1 << nw) ** 2, dtype=np.complex128)
    k = lib().orc_gate_matrix(name.encode(), nw, p, int(inverse), out)
    if k < 0:
        raise ValueError(f"unknown gate {name}/{nw}")
    return out.reshape(1 << nw, 1 << nw)


def generator(name: str, nw: int):
    out = np.zeros((1 << nw) ** 2, dtype=np.complex128)
    s = C.c_double()
    k = lib().orc_generator(name.encode(), nw, out, C.byref(s))
    if k < 0:
        raise ValueError(f"no generator for {name}/{nw}")
    return out.reshape(1 << nw, 1 << nw), s.value


def _pack_ops(ops):
    names = (C.c_char_p * max(1, len(ops)))()
    woff = [0]
    wflat, inv, pars, npar = [], [], [], []
    mats = (C.c_void_p * max(1, len(ops)))()
    keep = []
    for i, op in enumerate(ops):
        name, wires, params, inverse = op[:4]
        mat = op[4] if len(op) > 4 else None
        names[i] = name.encode()
        wflat.extend(int(w) for w in wires)
        woff.append(len(wflat))
        inv.append(int(bool(inverse)))
        p = list(params)[:3] + [0.0] * (3 - min(3, len(params)))
        pars.extend(float(x) for x in p)
        npar.append(len(params))
        if mat is not None:
            m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128).ravel())
            keep.append(m)
            mats[i] = m.ctypes.data
        else:
            mats[i] = None
    return (names, np.array(wflat or [0], dtype=np.int32), np.array(woff, dtype=np.int32),
            np.array(inv or [0], dtype=np.int32), np.array(pars or [0.0], dtype=np.float64),
            np.array(npar or [0], dtype=np.int32), mats, keep)


def apply_ops(n: int, psi: np.ndarray, ops) -> np.ndarray:
    """Apply ops to a copy of psi (complex128) and return it."""
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    if not ops:
        return out
    names, wflat, woff, inv, pars, _, mats, keep = _pack_ops(ops)
    _check(lib().orc_apply_ops(n, out, len(ops), names, wflat, woff, inv, pars, mats), "apply_ops")
    return out


def apply_generator(n: int, psi: np.ndarray, name: str, wires) -> np.ndarray:
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    w = np.array(wires, dtype=np.int32)
    _check(lib().orc_apply_generator(n, out, name.encode(), w, len(w)), "apply_generator")
    return out


def expval_matrix(n, psi, wires, M) -> complex:
    re, im = C.c_double(), C.c_double()
    w = np.array(wires, dtype=np.int32)
    m = np.ascontiguousarray(np.asarray(M, dtype=np.complex128).ravel())
    _check(lib().orc_expval_matrix(n, np.ascontiguousarray(psi, dtype=np.complex128), w, len(w), m,
                                   C.byref(re), C.byref(im)), "expval_matrix")
    return complex(re.value, im.value)


def expval_pauli_words(n, psi, words, wires, coeffs):
    """words: list of strings over IXYZ; wires: list of wire lists."""
    paulis = "".join(words).encode()
    wflat = np.array([w for ws in wires for w in ws] or [0], dtype=np.int32)
    offs = np.zeros(len(words) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(w) for w in words])
    per = np.zeros(max(1, len(words)))
    tot = C.c_double()
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128))
    _check(lib().orc_expval_pauli_words(n, np.ascontiguousarray(psi, dtype=np.complex128), len(words),
                                        paulis, wflat, offs, c, per, C.byref(tot)), "pauli_words")
    return tot.value, per[: len(words)]


def spmv(indptr, indices, data, x):
    y = np.zeros(len(indptr) - 1, dtype=np.complex128)
    _check(lib().orc_spmv(len(indptr) - 1, np.asarray(indptr, np.int64), np.asarray(indices, np.int64),
                          np.ascontiguousarray(data, np.complex128),
                          np.ascontiguousarray(x, np.complex128), y), "spmv")
    return y


def expval_csr(n, psi, indptr, indices, data) -> complex:
    re, im = C.c_double(), C.c_double()
    _check(lib().orc_expval_csr(n, np.ascontiguousarray(psi, np.complex128), np.asarray(indptr, np.int64),
                                np.asarray(indices, np.int64), np.ascontiguousarray(data, np.complex128),
                                C.byref(re), C.byref(im)), "expval_csr")
    return complex(re.value, im.value)


def probs(n, psi, wires) -> np.ndarray:
    """Marginal probabilities in the reference binding's order (wires[0] = LSB)."""
    w = np.array(wires, dtype=np.int32)
    out = np.zeros(1 << len(w))
    _check(lib().orc_probs(n, np.ascontiguousarray(psi, np.complex128), w, len(w), out), "probs")
    return out


def philox(seed: int, counter: int) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(seed, counter, out)
    return out


def sample(n, psi, shots, seed) -> np.ndarray:
    """Sorted basis indices drawn with the seeded fixed-point sampler."""
    out = np.zeros(max(1, shots), dtype=np.uint64)
    psi = np.ascontiguousarray(psi)
    if psi.dtype == np.complex64:
        _check(lib().orc_sample_c64(n, psi, shots, seed, out), "sample")
    else:
        _check(lib().orc_sample(n, psi.astype(np.complex128), shots, seed, out), "sample")
    return out[:shots]


def adjoint_jacobian(n, psi, bras, ops, trainable) -> np.ndarray:
    """Jacobian rows = observables (bras[j] = O_j psi), cols = trainable params."""
    names, wflat, woff, inv, pars, npar, mats, keep = _pack_ops(ops)
    bras = [np.ascontiguousarray(b, np.complex128) for b in bras]
    bp = (C.c_void_p * max(1, len(bras)))(*[b.ctypes.data for b in bras])
    tp = np.array(list(trainable) or [0], dtype=np.int64)
    jac = np.zeros(max(1, len(bras) * len(trainable)))
    _check(lib().orc_adjoint_jacobian(n, np.ascontiguousarray(psi, np.complex128), len(bras), bp,
                                      len(ops), names, wflat, woff, inv, pars, npar, mats, tp,
                                      len(trainable), jac), "adjoint_jacobian")
    return jac[: len(bras) * len(trainable)].reshape(len(bras), len(trainable))


def philox_raw(ctr, key) -> np.ndarray:
    L = lib()
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    L.orc_philox4x32_10_raw.restype = None
    L.orc_philox4x32_10_raw(c.ctypes.data_as(C.c_void_p), k.ctypes.data_as(C.c_void_p),
                            out.ctypes.data_as(C.c_void_p))
    return out
