"""Parity at the benchmark configurations' own sizes.

- 28 qubits complex128, 6 layers RY/RZ + CZ brick (498 ops, 336 parameters),
  <Z0 Z27>: the adjoint Jacobian through ``AdjointJacobianGPU_C128`` (the
  reference's AdjointJacobianGPU::adjointJacobian, AdjointDiffGPU.hpp:499-596)
  at its real launch shape -- the fused adjoint passes then step through
  several tiles per wavefront -- against the exact parameter-shift rule
  d<O>/dtheta = (<O>(theta + pi/2) - <O>(theta - pi/2)) / 2, which holds for
  RY and RZ (generator eigenvalues +-1/2), with both sides of the shift
  evaluated by GPU forward passes.
- 24 qubits: the 760-op RX/RY/RZ + CNOT-ladder state against the oracle, the
  256-word Pauli Hamiltonian, and 10^6 seeded samples (bit-exact).
- 20 qubits: the open-chain TFIM (J = 1, h = 0.7) as CSR, nnz = 21 * 2^20.

Tolerances: complex128 |d| <= 1e-10 * max(|ref|, 1e-3 * ||ref||_inf)
(``close128``); complex64 within 1e-5 of the complex128 reference, scaled
by max(1, ||ref||_inf) (``close64``); samples bit-exact.
"""
import math
import sys

import numpy as np
import pytest

from _tol import close128, close64
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

N = pytest.importorskip("pennylane_lightning_gpu._native")
ops_mod = pytest.importorskip("pennylane_lightning_gpu.lightning_gpu_qubit_ops")

from pennylane_lightning_gpu.workloads import ladder, random_pauli_hamiltonian, ry_rz_cz_brick, tfim_csr  # noqa: E402


def note(msg):
    print(msg, file=sys.stderr, flush=True)


def zz_expval(state, n):
    tot, _ = state.expval_pauli_words(["ZZ"], [[0, n - 1]], [1.0])
    return tot


@pytest.mark.parametrize("prefix", ["C128", "C64"])
def test_config2_28q_adjoint_vs_parameter_shift(prefix):
    n, layers, seed = 28, 6, 1337
    ops = ry_rz_cz_brick(n, layers, seed)
    assert len(ops) == 498
    npar = sum(1 for o in ops if o[2])
    assert npar == 336
    note(f"[fullsize] 28q adjoint {prefix}: forward + 336-column Jacobian")
    cls = getattr(ops_mod, f"LightningGPU_{prefix}")
    dt = np.float64 if prefix == "C128" else np.float32
    cdt = np.complex128 if prefix == "C128" else np.complex64
    sv = cls(n)
    for name, wires, params, inv in ops:
        getattr(sv, name)(wires, inv, params)
    adj = getattr(ops_mod, f"AdjointJacobianGPU_{prefix}")()
    obs = getattr(ops_mod, f"TensorProdObsGPU_{prefix}")(
        [getattr(ops_mod, f"NamedObsGPU_{prefix}")("PauliZ", [0]),
         getattr(ops_mod, f"NamedObsGPU_{prefix}")("PauliZ", [n - 1])])
    opsl = adj.create_ops_list([o[0] for o in ops], [np.array(o[2], dtype=dt) for o in ops],
                               [o[1] for o in ops], [o[3] for o in ops], [np.array([], dtype=cdt) for _ in ops])
    jac = np.asarray(adj.adjoint_jacobian(sv, [obs], opsl, list(range(npar))))[0]
    del sv
    assert jac.shape == (npar,) and np.all(np.isfinite(jac))
    # RY and RZ columns spread over all six layers, inside and outside the
    # light cone of Z0 Z27 (param index 2 * (28 * layer + wire) + {0: RY, 1: RZ})
    picks = [(0, 0, 0), (0, 27, 1), (1, 13, 0), (1, 1, 1), (2, 5, 1), (2, 26, 0), (3, 20, 0), (3, 3, 1),
             (4, 27, 0), (4, 14, 1), (5, 0, 1), (5, 27, 0)]
    cols = [2 * (n * layer + w) + k for layer, w, k in picks]
    pidx = [i for i, o in enumerate(ops) if o[2]]
    state = N.State(n)  # complex128 forward passes for the shift rule
    shift = []
    for c in cols:
        vals = []
        for sgn in (+1, -1):
            shifted = list(ops)
            name, wires, params, inv = shifted[pidx[c]]
            shifted[pidx[c]] = (name, wires, [params[0] + sgn * math.pi / 2], inv)
            state.reset().apply(shifted)
            vals.append(zz_expval(state, n))
        shift.append(0.5 * (vals[0] - vals[1]))
    shift = np.array(shift)
    note(f"[fullsize] 28q {prefix}: |J|_inf {np.max(np.abs(jac)):.3e}, shift columns {np.round(shift, 6)}")
    if prefix == "C128":
        # near-zero entries are held to 1e-10 of the Jacobian's largest entry
        scale = np.maximum(np.abs(shift), 1e-3 * np.max(np.abs(jac)))
        err = np.abs(jac[cols] - shift)
        assert np.all(err <= 1e-10 * scale), f"max rel err {np.max(err / scale):.3e}"
    else:
        close64(jac[cols], shift)


def config3_state(n, layers=8, seed=1337):
    ops = ladder(n, layers, seed)
    s = N.State(n).reset().apply(ops)
    return ops, s


def test_config3_24q_state_and_pauli_hamiltonian():
    n = 24
    ops, s = config3_state(n)
    assert len(ops) == 760
    note("[fullsize] 24q ladder state against the oracle (760 ops)")
    got = s.download()
    ref = po.apply_ops(n, po.zero_state(n), ops)
    close128(got, ref)
    del ref
    words, wires, coeffs = random_pauli_hamiltonian(n, 256, 1337)
    note("[fullsize] 24q Hamiltonian of 256 Pauli words")
    tot, per = s.expval_pauli_words(words, wires, coeffs)
    rt, rp = po.expval_pauli_words(n, got, words, wires, coeffs)
    close128(per, rp)
    close128(tot, rt)
    # complex64 on the same amplitudes
    s64 = N.State(n, np.complex64).upload(got.astype(np.complex64))
    t64, p64 = s64.expval_pauli_words(words, wires, coeffs)
    close64(p64, rp)
    close64(t64, rt)


def test_config3_24q_million_shots_bit_exact():
    n = 24
    _, s = config3_state(n)
    note("[fullsize] 24q, 10^6 seeded shots")
    got = s.sample(1_000_000, 1337)
    ref = po.sample(n, s.download(), 1_000_000, 1337)
    np.testing.assert_array_equal(got, ref)
    assert np.all(np.diff(got.astype(np.int64)) >= 0)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_config3_tfim20_csr(dtype):
    n = 20
    ops = ladder(n, 8, 1337)
    s128 = N.State(n).reset().apply(ops)
    psi = s128.download()
    close128(psi, po.apply_ops(n, po.zero_state(n), ops))
    indptr, indices, data = tfim_csr(n, 1.0, 0.7)
    assert len(data) == 21 * (1 << n)
    note(f"[fullsize] TFIM-20 CSR expval {np.dtype(dtype).name}")
    ref = po.expval_csr(n, psi, indptr, indices, data)
    s = s128 if dtype == np.complex128 else N.State(n, dtype).upload(psi.astype(dtype))
    c = N.CSR(indptr, indices, data, dtype=dtype)
    got = c.expval(s)
    (close128 if dtype == np.complex128 else close64)(got.real, ref.real)
    assert abs(ref.imag) < 1e-9


def test_single_gate_27q_workgroup_orders():
    """The 1-target pair kernel on a 2 GiB state, for targets that take the
    XCD-interleaved workgroup order (pair distance at byte-address bit 25)
    and the per-XCD eighths (bit 26), against numpy on the same seeded
    state; complex128, bit-for-bit up to the 1e-10 tolerance."""
    n = 27
    rng = np.random.default_rng(2727)
    psi = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)).astype(np.complex128)
    psi /= np.linalg.norm(psi)
    s = N.State(n).upload(psi)
    th = 0.731
    c, si = math.cos(th / 2), math.sin(th / 2)
    for b in (21, 22):
        note(f"[fullsize] 27q RX on amplitude bit {b}")
        s.apply([("RX", [n - 1 - b], [th], False)])
        v = psi.reshape(1 << (n - 1 - b), 2, 1 << b)
        a0, a1 = v[:, 0, :].copy(), v[:, 1, :].copy()
        v[:, 0, :] = c * a0 - 1j * si * a1
        v[:, 1, :] = -1j * si * a0 + c * a1
        del a0, a1
    got = s.download()
    close128(got, psi)
