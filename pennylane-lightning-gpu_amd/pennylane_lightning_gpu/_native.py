"""ctypes view of the C ABI (``include/lgh.h``) exported by ``liblgh.so``.

The pybind11 module ``lightning_gpu_qubit_ops`` is the drop-in boundary the
device class uses; this module gives Python direct access to the same C ABI
for callers that want it (the batching and serving forms, the torch autograd
function, and the C-ABI tests).  There is no fallback: if ``liblgh.so`` is
missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblgh.so")

LGH_C64 = 0
LGH_C128 = 1

GATES = [
    "Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "CNOT", "SWAP", "CSWAP",
    "Toffoli", "CY", "CZ", "PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "Rot", "CRX",
    "CRY", "CRZ", "CRot", "IsingXX", "IsingYY", "IsingZZ", "SingleExcitation",
    "SingleExcitationMinus", "SingleExcitationPlus", "DoubleExcitation", "DoubleExcitationMinus",
    "DoubleExcitationPlus", "MultiRZ", "Matrix",
]
GATE_ID = {name: i for i, name in enumerate(GATES)}
GATE_STATEPREP = -1
PARAM_COUNT = {"Rot": 3, "CRot": 3}
for _g in ("PhaseShift", "ControlledPhaseShift", "RX", "RY", "RZ", "CRX", "CRY", "CRZ", "IsingXX",
           "IsingYY", "IsingZZ", "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
           "DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ"):
    PARAM_COUNT[_g] = 1


class LghError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lgh error {code}: {msg}")
        self.code = code


class Op(C.Structure):
    _fields_ = [
        ("gate", C.c_int32),
        ("num_wires", C.c_int32),
        ("num_ctrl", C.c_int32),
        ("inverse", C.c_int32),
        ("wires", C.POINTER(C.c_int32)),
        ("params", C.c_double * 3),
        ("matrix", C.POINTER(C.c_double)),
        ("num_params", C.c_int32),
        ("reserved", C.c_int32),
    ]


class TuneResult(C.Structure):
    _fields_ = [("forward_tile_bits", C.c_int32), ("forward_waves_per_cu", C.c_int32),
                ("adjoint_tile_bits", C.c_int32), ("adjoint_waves_per_cu", C.c_int32),
                ("forward_ms", C.c_double), ("adjoint_ms", C.c_double),
                ("default_forward_ms", C.c_double), ("default_adjoint_ms", C.c_double),
                ("candidates", C.c_int32), ("cand_dir", C.c_int32 * 16), ("cand_tile_bits", C.c_int32 * 16),
                ("cand_waves_per_cu", C.c_int32 * 16), ("cand_ms", C.c_double * 16),
                ("adjoint_slots", C.c_int32), ("cand_slots", C.c_int32 * 16)]


_lib = None


class Request(C.Structure):
    _fields_ = [("num_qubits", C.c_int32), ("ops", C.POINTER(Op)), ("num_ops", C.c_int32),
                ("obs_paulis", C.POINTER(C.c_int32)), ("obs_wires", C.POINTER(C.c_int32)), ("obs_len", C.c_int32),
                ("want_grad", C.c_int32)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with build_ext.py")
        L = C.CDLL(LIB_PATH)
        L.lgh_last_error.restype = C.c_char_p
        vp = C.c_void_p
        L.lgh_sv_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, vp]
        L.lgh_sv_destroy.argtypes = [vp]
        L.lgh_sv_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(vp), C.POINTER(vp)]
        L.lgh_sv_reset.argtypes = [vp]
        L.lgh_sv_set_basis.argtypes = [vp, C.c_uint64, C.c_double, C.c_double]
        L.lgh_sv_set_indexed.argtypes = [vp, vp, vp, C.c_uint64]
        L.lgh_sv_copy_h2d.argtypes = [vp, vp, C.c_uint64]
        L.lgh_sv_copy_d2h.argtypes = [vp, vp, C.c_uint64]
        L.lgh_sv_copy_d2d.argtypes = [vp, vp]
        L.lgh_sv_synchronize.argtypes = [vp]
        L.lgh_apply.argtypes = [vp, C.POINTER(Op), C.c_int64]
        L.lgh_gate_matrix.argtypes = [C.POINTER(Op), vp]
        L.lgh_apply_generator.argtypes = [vp, C.POINTER(Op), C.POINTER(C.c_double)]
        L.lgh_inner.argtypes = [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.lgh_axpy.argtypes = [vp, C.c_double, C.c_double, vp]
        L.lgh_sv_pauli_word_from.argtypes = [vp, vp, C.c_char_p, vp, C.c_int32]
        L.lgh_scale.argtypes = [vp, C.c_double, C.c_double]
        L.lgh_expval_matrix.argtypes = [vp, vp, C.c_int32, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.lgh_expval_pauli_words.argtypes = [vp, C.c_int64, C.c_char_p, vp, vp, vp, vp, C.POINTER(C.c_double)]
        L.lgh_pauli_plan.argtypes = [C.c_int32, C.c_int64, C.c_char_p, vp, vp] + [C.POINTER(C.c_int32)] * 3
        L.lgh_expval_pauli_words_noalloc.argtypes = [vp, C.c_int64, C.c_char_p, vp, vp, vp, vp,
                                                     C.POINTER(C.c_double)]
        L.lgh_csr_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int64, vp, vp, vp, C.c_int64]
        L.lgh_csr_destroy.argtypes = [vp]
        L.lgh_expval_csr.argtypes = [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.lgh_apply_csr.argtypes = [vp, vp, vp]
        L.lgh_probs.argtypes = [vp, vp, C.c_int32, vp]
        L.lgh_sample.argtypes = [vp, C.c_uint64, C.c_uint64, vp]
        L.lgh_sample_bits.argtypes = [vp, C.c_uint64, C.c_uint64, vp]
        L.lgh_adjoint_jacobian.argtypes = [vp, C.POINTER(vp), C.c_int32, C.POINTER(Op), C.c_int64, vp,
                                           C.c_int64, C.c_int32, vp]
        L.lgh_adjoint_jacobian_zwords.argtypes = [vp, C.POINTER(vp), vp, C.c_int32, C.POINTER(Op), C.c_int64, vp,
                                                  C.c_int64, C.c_int32, vp]
        L.lgh_device_count.argtypes = [C.POINTER(C.c_int)]
        L.lgh_device_arch.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
        L.lgh_device_supported.argtypes = [C.c_int, C.POINTER(C.c_int)]
        L.lgh_device_memory.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.lgh_device_synchronize.argtypes = [C.c_int]
        L.lgh_pool_trim.argtypes = [C.c_int]
        L.lgh_stream_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.lgh_stream_destroy.argtypes = [vp]
        L.lgh_stream_synchronize.argtypes = [vp]
        L.lgh_set_tile_bits.argtypes = [C.c_int, C.c_int]
        L.lgh_set_pass_waves.argtypes = [C.c_int64, C.c_int64]
        L.lgh_set_lean_passes.argtypes = [C.c_int]
        L.lgh_autotune_fused.argtypes = [C.c_int, C.c_int32, C.c_int, vp, C.c_int64, C.POINTER(TuneResult)]
        L.lgh_fusion_plan.argtypes = [C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_int] + [C.POINTER(C.c_int32)] * 4
        L.lgh_fusion_lean_passes.argtypes = [C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_int,
                                             C.POINTER(C.c_int32)]
        L.lgh_profile_enable.argtypes = [C.c_int]
        L.lgh_profile_read.argtypes = [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.lgh_tmpl_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(Op), C.c_int32, vp, vp, vp, C.c_int32,
                                      C.c_int32]
        L.lgh_tmpl_destroy.argtypes = [vp]
        L.lgh_tmpl_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.lgh_tmpl_passes.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.lgh_batch_run.argtypes = [vp, C.c_int32, C.c_int, vp, vp, vp, vp, vp, vp]
        L.lgh_batch_run_serial.argtypes = [vp, C.c_int32, C.c_int, vp, vp, vp, vp, vp, vp]
        L.lgh_batch_plan_create.argtypes = [C.POINTER(vp), vp, C.c_int32, C.c_int, C.c_int32]
        L.lgh_batch_plan_run.argtypes = [vp, vp, vp, vp, vp, vp]
        L.lgh_batch_plan_info.argtypes = [vp, C.POINTER(C.c_int32), vp]
        L.lgh_batch_plan_destroy.argtypes = [vp]
        L.lgh_server_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_uint64, C.c_int32, C.c_int32]
        L.lgh_server_destroy.argtypes = [vp]
        L.lgh_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.lgh_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(C.c_double), vp, C.c_int64, vp]
        L.lgh_server_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]
        L.lgh_server_set_limits.argtypes = [vp, C.c_int32, C.c_int32]
        L.lgh_server_inject_failures.argtypes = [vp, C.c_uint64]
        L.lgh_server_set_batching.argtypes = [vp, C.c_int32, C.c_int32]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise LghError(rc, lib().lgh_last_error().decode())


def set_fusion(enabled: bool):
    """Switch gate fusion of lgh_apply sequences and the adjoint sweep."""
    check(lib().lgh_set_fusion(int(bool(enabled))))


def set_tile_bits(forward: int, adjoint: int):
    """Register bits (3 or 4) of the fused forward and adjoint passes
    (adjoint 0: the default, 4 for complex128 states of 24 qubits and more,
    else 3)."""
    check(lib().lgh_set_tile_bits(int(forward), int(adjoint)))


def set_pass_waves(forward: int, adjoint: int):
    """Caps on the wavefronts of the fused passes' grids (0: the defaults);
    each wavefront then steps through a run of consecutive tiles."""
    check(lib().lgh_set_pass_waves(int(forward), int(adjoint)))


def pauli_plan(num_qubits, words, wires):
    """Host-only plan of a Pauli-word expectation: dict with ``tile_passes``,
    ``tile_launches`` and ``pair_launches`` (lgh_pauli_plan)."""
    chars = "".join(words).encode()
    wf = np.ascontiguousarray([x for ws in wires for x in ws] or [0], dtype=np.int32)
    offs = np.zeros(len(words) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(w) for w in words])
    v = [C.c_int32() for _ in range(3)]
    check(lib().lgh_pauli_plan(int(num_qubits), len(words), chars, wf.ctypes.data, offs.ctypes.data,
                               *[C.byref(x) for x in v]))
    return {"tile_passes": v[0].value, "tile_launches": v[1].value, "pair_launches": v[2].value}


def autotune_fused(num_qubits, ops=None, dtype=np.complex128, device=0):
    """Times the fused passes of ``ops`` (default: 2 layers RY/RZ + CZ brick)
    on ``num_qubits`` qubits under every candidate tile height and grid and
    keeps the fastest forward and adjoint choice for that size and dtype
    (lgh_autotune_fused).  Returns the choice and every candidate's time."""
    r = TuneResult()
    code = LGH_C128 if np.dtype(dtype) == np.complex128 else LGH_C64
    if ops:
        arr, keep = make_ops(ops)
        check(lib().lgh_autotune_fused(int(device), int(num_qubits), code, arr, len(ops), C.byref(r)))
    else:
        check(lib().lgh_autotune_fused(int(device), int(num_qubits), code, None, 0, C.byref(r)))
    cands = [{"dir": "forward" if r.cand_dir[q] == 0 else "adjoint", "tile_bits": r.cand_tile_bits[q],
              "waves_per_cu": r.cand_waves_per_cu[q], "ms": round(r.cand_ms[q], 4)} for q in range(r.candidates)]
    for q, c in enumerate(cands):
        if c["dir"] == "adjoint":
            c["slots"] = r.cand_slots[q]
    return {"forward": {"tile_bits": r.forward_tile_bits, "waves_per_cu": r.forward_waves_per_cu,
                        "ms": round(r.forward_ms, 4), "default_ms": round(r.default_forward_ms, 4)},
            "adjoint": {"tile_bits": r.adjoint_tile_bits, "waves_per_cu": r.adjoint_waves_per_cu,
                        "slots": r.adjoint_slots, "ms": round(r.adjoint_ms, 4),
                        "default_ms": round(r.default_adjoint_ms, 4)},
            "candidates": cands}


def autotune_clear():
    """Drops every tuned choice (the defaults apply again)."""
    check(lib().lgh_autotune_clear())


def fusion_plan(n, ops, adjoint=False, dtype=np.complex128):
    """Fused-pass plan of ``ops`` on ``n`` qubits, computed on the host (no
    device): dict with ``passes``, ``standalone``, ``runs`` (diagonal runs)
    and ``run_members`` (diagonal gates applied inside runs)."""
    arr, keep = make_ops(ops) if ops else (None, None)
    out = [C.c_int32() for _ in range(4)]
    code = LGH_C128 if np.dtype(dtype) == np.complex128 else LGH_C64
    check(lib().lgh_fusion_plan(int(n), arr, len(ops), int(bool(adjoint)), code, *[C.byref(o) for o in out]))
    return dict(zip(("passes", "standalone", "runs", "run_members"), (o.value for o in out)))


def fusion_lean_passes(n, ops, adjoint=False, dtype=np.complex128) -> int:
    """How many of the fused passes of ``ops`` run in the lean kernels (host only)."""
    arr, keep = make_ops(ops) if ops else (None, None)
    out = C.c_int32()
    code = LGH_C128 if np.dtype(dtype) == np.complex128 else LGH_C64
    check(lib().lgh_fusion_lean_passes(int(n), arr, len(ops), int(bool(adjoint)), code, C.byref(out)))
    return out.value


def device_count() -> int:
    c = C.c_int()
    check(lib().lgh_device_count(C.byref(c)))
    return c.value


# numpy view of struct lgh_op (same field order, offsets and size)
OP_DTYPE = np.dtype([("gate", np.int32), ("num_wires", np.int32), ("num_ctrl", np.int32), ("inverse", np.int32),
                     ("wires", np.uint64), ("params", np.float64, (3,)), ("matrix", np.uint64),
                     ("num_params", np.int32), ("reserved", np.int32)])
assert OP_DTYPE.itemsize == C.sizeof(Op)
_STATEPREP_NAMES = ("StatePrep", "QubitStateVector", "BasisState")


def make_ops(ops):
    """Pack ``(name, wires, params, inverse[, matrix])`` tuples into an lgh_op array.

    Returns the array (a ctypes view of a numpy record array) and a list of
    buffers that must outlive the call.  Wires of all ops share one int32
    array; the records are filled column by column.
    """
    n = len(ops)
    rec = np.zeros(max(1, n), dtype=OP_DTYPE)
    keep = [rec]
    if n == 0:
        return (Op * 1).from_buffer(rec), keep
    gate = np.empty(n, dtype=np.int32)
    nwires = np.empty(n, dtype=np.int64)
    params = np.zeros((n, 3))
    npar = np.empty(n, dtype=np.int32)
    inv = np.empty(n, dtype=np.int32)
    flat = []
    mat_id = GATE_ID["Matrix"]
    for i, op in enumerate(ops):
        name, wires, ps, inverse = op[:4]
        g = GATE_ID.get(name)
        if g is None or g == mat_id:
            if name in _STATEPREP_NAMES:
                g = GATE_STATEPREP
            elif len(op) > 4 and op[4] is not None:
                g = mat_id
                m = np.ascontiguousarray(np.asarray(op[4], dtype=np.complex128).ravel()).view(np.float64)
                keep.append(m)
                rec["matrix"][i] = m.ctypes.data
            else:
                raise ValueError(f"Currently unsupported gate: {name}")
        gate[i] = g
        nwires[i] = len(wires)
        flat.extend(wires)
        k = len(ps)
        npar[i] = k
        if k:
            params[i, :min(k, 3)] = list(ps)[:3]
        inv[i] = 1 if inverse else 0
    wires_arr = np.asarray(flat if flat else [0], dtype=np.int32)
    keep.append(wires_arr)
    starts = np.zeros(n, dtype=np.int64)
    starts[1:] = np.cumsum(nwires[:-1])
    rec["gate"][:n] = gate
    rec["num_wires"][:n] = nwires
    rec["inverse"][:n] = inv
    rec["wires"][:n] = np.uint64(wires_arr.ctypes.data) + starts.astype(np.uint64) * np.uint64(4)
    rec["params"][:n] = params
    rec["num_params"][:n] = npar
    return (Op * n).from_buffer(rec), keep


def gate_matrix(name, wires, params=(), inverse=False, matrix=None):
    """Host-side matrix the library applies for one op (no device needed)."""
    op = (name, wires, params, inverse) + ((matrix,) if matrix is not None else ())
    arr, keep = make_ops([op])
    dim = 1 << len(wires)
    out = np.zeros(dim * dim, dtype=np.complex128)
    check(lib().lgh_gate_matrix(arr, out.ctypes.data))
    return out.reshape(dim, dim)


class State:
    """Owning handle of one device state (``lgh_sv``)."""

    def __init__(self, num_qubits: int, dtype=np.complex128, device: int = 0, stream=None):
        self.n = int(num_qubits)
        self.dtype = np.dtype(dtype)
        code = LGH_C128 if self.dtype == np.complex128 else LGH_C64
        h = C.c_void_p()
        check(lib().lgh_sv_create(C.byref(h), self.n, code, device, stream))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.lgh_sv_destroy(h)
            self.handle = None

    def close(self):
        self.__del__()

    @property
    def length(self):
        return 1 << self.n

    def data_ptr(self) -> int:
        p = C.c_void_p()
        check(lib().lgh_sv_info(self.handle, None, None, None, None, C.byref(p)))
        return p.value

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().lgh_sv_info(self.handle, None, None, None, C.byref(s), None))
        return s.value or 0

    def reset(self):
        check(lib().lgh_sv_reset(self.handle))
        return self

    def set_basis(self, index: int, value=1.0):
        v = complex(value)
        check(lib().lgh_sv_set_basis(self.handle, index, v.real, v.imag))
        return self

    def set_indexed(self, indices, values):
        idx = np.ascontiguousarray(indices, dtype=np.int64)
        val = np.ascontiguousarray(values, dtype=self.dtype)
        check(lib().lgh_sv_set_indexed(self.handle, idx.ctypes.data, val.ctypes.data, len(idx)))
        return self

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        check(lib().lgh_sv_copy_h2d(self.handle, h.ctypes.data, h.size))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.length, dtype=self.dtype)
        check(lib().lgh_sv_copy_d2h(self.handle, out.ctypes.data, out.size))
        return out

    def copy_from(self, other: "State"):
        check(lib().lgh_sv_copy_d2d(self.handle, other.handle))
        return self

    def synchronize(self):
        check(lib().lgh_sv_synchronize(self.handle))

    def apply(self, ops):
        if not ops:
            return self
        arr, keep = make_ops(ops)
        check(lib().lgh_apply(self.handle, arr, len(ops)))
        return self

    def apply_generator(self, name, wires):
        arr, keep = make_ops([(name, wires, [0.0] * PARAM_COUNT.get(name, 0), False)])
        s = C.c_double()
        check(lib().lgh_apply_generator(self.handle, arr, C.byref(s)))
        return s.value

    def inner(self, other: "State") -> complex:
        re, im = C.c_double(), C.c_double()
        check(lib().lgh_inner(self.handle, other.handle, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def axpy(self, alpha, x: "State"):
        a = complex(alpha)
        check(lib().lgh_axpy(self.handle, a.real, a.imag, x.handle))

    def set_pauli_word_of(self, src: "State", word: str, wires):
        """self <- P src for the Pauli word `word` (letters I/X/Y/Z) on `wires`; one pass."""
        w = np.ascontiguousarray(wires, dtype=np.int32)
        if len(word) != len(w):
            raise ValueError("one wire per Pauli letter")
        check(lib().lgh_sv_pauli_word_from(self.handle, src.handle, word.encode(), w.ctypes.data, len(w)))
        return self

    def scale(self, alpha):
        a = complex(alpha)
        check(lib().lgh_scale(self.handle, a.real, a.imag))

    def expval_matrix(self, wires, M) -> complex:
        w = np.ascontiguousarray(wires, dtype=np.int32)
        m = np.ascontiguousarray(np.asarray(M, dtype=np.complex128).ravel())
        re, im = C.c_double(), C.c_double()
        check(lib().lgh_expval_matrix(self.handle, w.ctypes.data, len(w), m.ctypes.data, C.byref(re),
                                      C.byref(im)))
        return complex(re.value, im.value)

    def expval_pauli_words(self, words, wires, coeffs, noalloc=False):
        """sum_w Re(c_w <P_w>) and the per-word <P_w>.  noalloc: through
        lgh_expval_pauli_words_noalloc (fixed device scratch, at most 1024
        words, no allocation after the first call on a device)."""
        chars = "".join(words).encode()
        wf = np.ascontiguousarray([x for ws in wires for x in ws] or [0], dtype=np.int32)
        offs = np.zeros(len(words) + 1, dtype=np.int64)
        offs[1:] = np.cumsum([len(w) for w in words])
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128))
        per = np.zeros(max(1, len(words)))
        tot = C.c_double()
        fn = lib().lgh_expval_pauli_words_noalloc if noalloc else lib().lgh_expval_pauli_words
        check(fn(self.handle, len(words), chars, wf.ctypes.data, offs.ctypes.data, c.ctypes.data, per.ctypes.data,
                 C.byref(tot)))
        return tot.value, per[: len(words)]

    def probs(self, wires) -> np.ndarray:
        w = np.ascontiguousarray(wires, dtype=np.int32)
        out = np.zeros(1 << len(w))
        check(lib().lgh_probs(self.handle, w.ctypes.data, len(w), out.ctypes.data))
        return out

    def sample(self, shots, seed) -> np.ndarray:
        out = np.zeros(max(1, shots), dtype=np.uint64)
        check(lib().lgh_sample(self.handle, shots, seed, out.ctypes.data))
        return out[:shots]

    def sample_bits(self, shots, seed) -> np.ndarray:
        out = np.zeros((max(1, shots), self.n), dtype=np.uint64)
        check(lib().lgh_sample_bits(self.handle, shots, seed, out.ctypes.data))
        return out[:shots]


class CSR:
    """Device-resident CSR matrix (``lgh_csr``)."""

    def __init__(self, indptr, indices, data, dtype=np.complex128, device=0):
        self.indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.int64)
        vals = np.ascontiguousarray(np.asarray(data, dtype=np.complex128))
        h = C.c_void_p()
        code = LGH_C128 if np.dtype(dtype) == np.complex128 else LGH_C64
        check(lib().lgh_csr_create(C.byref(h), device, code, len(self.indptr) - 1, self.indptr.ctypes.data,
                                   self.indices.ctypes.data, vals.ctypes.data, len(self.indices)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.lgh_csr_destroy(h)
            self.handle = None

    def expval(self, state: State) -> complex:
        re, im = C.c_double(), C.c_double()
        check(lib().lgh_expval_csr(state.handle, self.handle, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def apply(self, out: State, x: State):
        check(lib().lgh_apply_csr(out.handle, self.handle, x.handle))


def adjoint_jacobian(psi: State, bras, ops, trainable, apply_operations=False) -> np.ndarray:
    """Jacobian [obs][trainable]; ``bras[j]`` holds O_j psi on entry and is consumed."""
    arr, keep = make_ops(ops)
    hs = (C.c_void_p * max(1, len(bras)))(*[b.handle.value for b in bras])
    tp = np.ascontiguousarray(list(trainable) or [0], dtype=np.int64)
    jac = np.zeros(max(1, len(bras) * len(trainable)))
    check(lib().lgh_adjoint_jacobian(psi.handle, hs, len(bras), arr, len(ops), tp.ctypes.data, len(trainable),
                                     int(apply_operations), jac.ctypes.data))
    return jac[: len(bras) * len(trainable)].reshape(len(bras), len(trainable))


def adjoint_jacobian_zwords(psi: State, zwires, ops, trainable, apply_operations=False) -> np.ndarray:
    """Jacobian [obs][trainable] for observables that are Z words: zwires[j]
    lists the wires of observable j's Z factors.  The observable states are
    scratch states formed by the sweep itself (lgh_adjoint_jacobian_zwords)."""
    arr, keep = make_ops(ops)
    bras = [State(psi.n, psi.dtype) for _ in zwires]
    hs = (C.c_void_p * max(1, len(bras)))(*[b.handle.value for b in bras])
    zm = np.ascontiguousarray([sum(1 << (psi.n - 1 - int(w)) for w in ws) for ws in zwires] or [0], dtype=np.uint64)
    tp = np.ascontiguousarray(list(trainable) or [0], dtype=np.int64)
    jac = np.zeros(max(1, len(bras) * len(trainable)))
    check(lib().lgh_adjoint_jacobian_zwords(psi.handle, hs, zm.ctypes.data, len(bras), arr, len(ops), tp.ctypes.data,
                                            len(trainable), int(apply_operations), jac.ctypes.data))
    return jac[: len(bras) * len(trainable)].reshape(len(bras), len(trainable))


# ---------------------------------------------------------------- launch timing

PROFILE_KINDS = {"fused_pass": 0, "fused_adjoint": 1, "batch": 2}


def set_lean_passes(enabled: bool):
    """Run passes of lean gate forms in the lean kernels (default on)."""
    check(lib().lgh_set_lean_passes(int(bool(enabled))))


def profile_enable(enabled: bool = True):
    """Record HIP events around every launch of the hot kernels (clears what was recorded)."""
    check(lib().lgh_profile_enable(int(bool(enabled))))


def profile_read(kind: str):
    """(launches, total milliseconds) recorded for ``kind`` since profiling was enabled."""
    n, t = C.c_int64(), C.c_double()
    check(lib().lgh_profile_read(PROFILE_KINDS[kind], C.byref(n), C.byref(t)))
    return n.value, t.value


# ---------------------------------------------------------------- batched circuits

PAULI_GATE = {"X": GATE_ID["PauliX"], "Y": GATE_ID["PauliY"], "Z": GATE_ID["PauliZ"]}


class Template:
    """Gate structure shared by circuits that differ only in their parameters.

    ``ops``: ``(name, wires, params, inverse)`` tuples; ``param_index[i]`` is the
    circuit parameter driving op i (``-1``: fixed as given).  ``observable``:
    Pauli word as ``[("Z", wire), ...]``.
    """

    def __init__(self, num_qubits, ops, param_index, observable, device=0):
        arr, keep = make_ops(ops)
        pidx = np.ascontiguousarray(list(param_index) or [0], dtype=np.int32)
        obs_p = np.ascontiguousarray([PAULI_GATE[p] for p, _ in observable] or [0], dtype=np.int32)
        obs_w = np.ascontiguousarray([int(w) for _, w in observable] or [0], dtype=np.int32)
        self.handle = C.c_void_p()
        check(lib().lgh_tmpl_create(C.byref(self.handle), int(num_qubits), arr, len(ops), pidx.ctypes.data,
                                    obs_p.ctypes.data, obs_w.ctypes.data, len(observable), int(device)))
        npar, nq = C.c_int32(), C.c_int32()
        check(lib().lgh_tmpl_info(self.handle, C.byref(npar), C.byref(nq)))
        self.num_params, self.num_qubits = npar.value, nq.value
        self._plans = weakref.WeakSet()  # batch plans that use this template

    def passes(self):
        """(forward passes, reverse passes, qubits of the states) of a batched run."""
        f, r, q = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().lgh_tmpl_passes(self.handle, C.byref(f), C.byref(r), C.byref(q)))
        return f.value, r.value, q.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def close(self):
        """Destroys the template; batch plans built on it are closed first
        (a plan must not outlive its templates)."""
        for plan in list(getattr(self, "_plans", ())):
            plan.close()
        if getattr(self, "handle", None) and self.handle.value:
            lib().lgh_tmpl_destroy(self.handle)
            self.handle = C.c_void_p()


def batch_run(templates, params, grad=True, dtype=np.complex128, serial=False, stream=None):
    """Run circuit c = ``templates[c]`` with ``params[c]`` (a 1-D array per circuit).

    Returns ``(expvals, jacobians)``: expvals[c] = <P>, jacobians[c] = d<P>/d params[c]
    (None without ``grad``).
    """
    nc = len(templates)
    hs = (C.c_void_p * max(1, nc))(*[t.handle.value for t in templates])
    flat = [np.asarray(p, dtype=np.float64).ravel() for p in params]
    for t, p in zip(templates, flat):
        if p.size != t.num_params:
            raise ValueError(f"template takes {t.num_params} parameters, got {p.size}")
    offs = np.zeros(max(1, nc), dtype=np.int64)
    if nc:
        offs[1:nc] = np.cumsum([p.size for p in flat])[:-1]
    allp = np.ascontiguousarray(np.concatenate(flat) if flat and sum(p.size for p in flat) else np.zeros(1))
    ev = np.zeros(max(1, nc))
    jac = np.zeros(max(1, int(sum(p.size for p in flat)))) if grad else None
    fn = lib().lgh_batch_run_serial if serial else lib().lgh_batch_run
    code = LGH_C128 if np.dtype(dtype) == np.complex128 else LGH_C64
    check(fn(hs, nc, code, allp.ctypes.data, offs.ctypes.data, ev.ctypes.data,
             jac.ctypes.data if grad else None, offs.ctypes.data if grad else None, stream))
    jacs = [jac[o:o + p.size].copy() for o, p in zip(offs[:nc], flat)] if grad else None
    return ev[:nc].copy(), jacs


class BatchPlan:
    """A batch of fixed composition prepared once (lgh_batch_plan_*): the
    launch sequence is a captured HIP graph replayed by every ``run``, and
    the persistent launch's grid is autotuned at creation."""

    def __init__(self, templates, grad=True, dtype=np.complex128):
        self.templates = list(templates)
        self.nc = len(self.templates)
        self.grad = bool(grad)
        self.handle = C.c_void_p()
        self._hs = (C.c_void_p * max(1, self.nc))(*[t.handle.value for t in self.templates])
        code = LGH_C128 if np.dtype(dtype) == np.complex128 else LGH_C64
        check(lib().lgh_batch_plan_create(C.byref(self.handle), self._hs, self.nc, code, int(self.grad)))
        for t in set(self.templates):
            t._plans.add(self)
        sizes = [t.num_params for t in self.templates]
        self.offs = np.zeros(max(1, self.nc), dtype=np.int64)
        if self.nc:
            self.offs[1:self.nc] = np.cumsum(sizes)[:-1]
        self.total = int(sum(sizes))

    def run(self, params):
        """params: (num_circuits, P) array or a list of 1-D arrays; returns
        (expvals, jacobians or None) like ``batch_run``."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in params])
                                    if self.total else np.zeros(1))
        if flat.size != max(1, self.total) and self.total:
            raise ValueError(f"plan takes {self.total} parameters in total, got {flat.size}")
        if not self.handle.value:
            raise LghError(1, "batch plan is closed")
        ev = np.zeros(max(1, self.nc))
        jac = np.zeros(max(1, self.total)) if self.grad else None
        check(lib().lgh_batch_plan_run(self.handle, flat.ctypes.data, self.offs.ctypes.data, ev.ctypes.data,
                                       jac.ctypes.data if self.grad else None,
                                       self.offs.ctypes.data if self.grad else None))
        sizes = [t.num_params for t in self.templates]
        jacs = [jac[o:o + k].copy() for o, k in zip(self.offs[:self.nc], sizes)] if self.grad else None
        return ev[:self.nc].copy(), jacs

    def info(self):
        g = C.c_int32()
        ms = np.zeros(2)
        check(lib().lgh_batch_plan_info(self.handle, C.byref(g), ms.ctypes.data))
        return {"grid_per_cu": g.value, "tuned_ms": ms.tolist()}

    def close(self):
        if getattr(self, "handle", None) and self.handle.value:
            check(lib().lgh_batch_plan_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving

class Server:
    """Request queue in front of one device (lgh_server_*): continuous
    batching of small circuits, one state per large circuit, memory admission.

    ``submit(n, ops, observable, grad)`` returns a ticket; ``wait(ticket)``
    returns ``(expval, grad or None, (queued_ms, run_ms, total_ms))``.
    """

    def __init__(self, device=0, memory_budget=0, num_streams=4, batch_max_qubits=20):
        self.handle = C.c_void_p()
        check(lib().lgh_server_create(C.byref(self.handle), int(device), int(memory_budget), int(num_streams),
                                      int(batch_max_qubits)))
        self._npar = {}

    @staticmethod
    def prepare(num_qubits, ops, observable, grad=False):
        """Packs a request once (ops, observable) for ``submit_prepared``: the
        packing is host work of the client, kept out of a submission loop."""
        arr, keep = make_ops(ops)
        obs_p = np.ascontiguousarray([PAULI_GATE[p] for p, _ in observable] or [0], dtype=np.int32)
        obs_w = np.ascontiguousarray([int(w) for _, w in observable] or [0], dtype=np.int32)
        req = Request(int(num_qubits), C.cast(arr, C.POINTER(Op)), len(ops),
                      obs_p.ctypes.data_as(C.POINTER(C.c_int32)), obs_w.ctypes.data_as(C.POINTER(C.c_int32)),
                      len(observable), int(bool(grad)))
        npar = sum(1 for o in ops if len(o[2]) > 0)
        return (req, (arr, keep, obs_p, obs_w), npar, bool(grad))

    def submit_prepared(self, prepared) -> int:
        req, _keep, npar, grad = prepared
        t = C.c_uint64()
        check(lib().lgh_server_submit(self.handle, C.byref(req), C.byref(t)))
        self._npar[t.value] = (npar, grad)
        return t.value

    def submit(self, num_qubits, ops, observable, grad=False) -> int:
        return self.submit_prepared(self.prepare(num_qubits, ops, observable, grad))

    def wait(self, ticket):
        npar, grad = self._npar.pop(ticket)
        e = C.c_double()
        g = np.zeros(max(1, npar))
        tm = np.zeros(3)
        check(lib().lgh_server_wait(self.handle, int(ticket), C.byref(e), g.ctypes.data, g.size, tm.ctypes.data))
        return e.value, (g[:npar].copy() if grad else None), tuple(tm)

    def set_limits(self, max_concurrent_batches, max_batch_circuits=4096):
        """Batches that may run at once, circuits per batch."""
        check(lib().lgh_server_set_limits(self.handle, int(max_concurrent_batches), int(max_batch_circuits)))

    def set_batching(self, min_batch, wait_us):
        """Under load, batches below min_batch circuits wait up to wait_us for more."""
        check(lib().lgh_server_set_batching(self.handle, int(min_batch), int(wait_us)))

    def inject_failures(self, count):
        """The next ``count`` pieces of work fail before launching (tests)."""
        check(lib().lgh_server_inject_failures(self.handle, int(count)))

    def stats(self):
        v = [C.c_uint64() for _ in range(4)]
        check(lib().lgh_server_stats(self.handle, *[C.byref(x) for x in v]))
        return {"batches": v[0].value, "batched_requests": v[1].value, "single_requests": v[2].value,
                "budget": v[3].value}

    def close(self):
        if getattr(self, "handle", None) and self.handle.value:
            check(lib().lgh_server_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
