"""Batched circuits (lgh_tmpl_create / lgh_batch_run) against the CPU oracle.

Each circuit is a template's gates with the circuit's own parameters; the
oracle runs the same gates on |0..0>, takes <P> and its adjoint Jacobian,
and sums the Jacobian columns of gates that share a parameter.  The serial
form (one state and one device call sequence per circuit) must agree with the
persistent batched kernel, and central finite differences (h = 1e-4, error
O(h^2) ~ 1e-8) pin the gradient independently of the oracle's adjoint.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from _tol import close128, close64

pytestmark = pytest.mark.gpu

N = pytest.importorskip("pennylane_lightning_gpu._native")


def ring_template(n, layers=3):
    """Per layer RX, RY, RZ on every wire, then a CNOT ring (the batched
    workload's circuit family): 12n gates, 9n parameters for 3 layers."""
    ops, pidx, k = [], [], 0
    for _ in range(layers):
        for w in range(n):
            for g in ("RX", "RY", "RZ"):
                ops.append((g, [w], [0.0], False))
                pidx.append(k)
                k += 1
        for w in range(n):
            ops.append(("CNOT", [w, (w + 1) % n], [], False))
            pidx.append(-1)
    return ops, pidx


def bind(ops, pidx, theta):
    return [(o[0], o[1], [theta[p]] if p >= 0 else o[2], o[3]) for o, p in zip(ops, pidx)]


def oracle_run(n, ops, pidx, theta, obs):
    """<P> and d<P>/d theta through the oracle."""
    bound = bind(ops, pidx, theta)
    psi = po.apply_ops(n, po.zero_state(n), bound)
    pops = [("Pauli" + p, [w], [], False) for p, w in obs]
    bra = po.apply_ops(n, psi, pops)
    e = float(np.real(np.vdot(psi, bra)))
    par_ops = [i for i, o in enumerate(bound) if o[2]]
    tr = [j for j, i in enumerate(par_ops) if pidx[i] >= 0]
    g = np.zeros(max(pidx) + 1 if max(pidx) >= 0 else 0)
    if not tr:
        return e, g
    jac_ops = po.adjoint_jacobian(n, psi, [bra], bound, tr)[0]
    for col, j in enumerate(tr):
        g[pidx[par_ops[j]]] += jac_ops[col]
    return e, g


@pytest.mark.parametrize("n", [3, 8, 10, 11, 13])
def test_batch_ring_matches_oracle(n):
    rng = np.random.default_rng(100 + n)
    ops, pidx = ring_template(n)
    obs = [("Z", 0)]
    t = N.Template(n, ops, pidx, obs)
    thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for _ in range(3)]
    ev, jac = N.batch_run([t] * 3, thetas)
    for c in range(3):
        e, g = oracle_run(n, ops, pidx, thetas[c], obs)
        close128(ev[c], e)
        close128(jac[c], g)


def test_batch_c64_matches_oracle():
    rng = np.random.default_rng(7)
    n = 12
    ops, pidx = ring_template(n, layers=2)
    obs = [("Z", 0), ("Z", n - 1)]
    t = N.Template(n, ops, pidx, obs)
    theta = rng.uniform(0, 2 * np.pi, t.num_params)
    ev, jac = N.batch_run([t], [theta], dtype=np.complex64)
    e, g = oracle_run(n, ops, pidx, theta, obs)
    close64(ev[0], e)
    close64(jac[0], g)


def test_batch_mixed_gates_observable_and_sharing():
    """Controlled and diagonal parametric gates, inverse flags, fixed gates,
    a shared parameter and an X/Y/Z observable."""
    rng = np.random.default_rng(11)
    n = 11
    ops = [("Hadamard", [0], [], False), ("RX", [1], [0.0], False), ("CRY", [1, 4], [0.0], False),
           ("CRZ", [4, 10], [0.0], True), ("PhaseShift", [2], [0.0], False), ("S", [3], [], False),
           ("ControlledPhaseShift", [2, 9], [0.0], False), ("IsingZZ", [5, 7], [0.0], False),
           ("MultiRZ", [0, 6, 8], [0.0], True), ("CRX", [8, 3], [0.0], False), ("Toffoli", [0, 1, 2], [], False),
           ("T", [9], [], True), ("RY", [10], [0.0], True), ("CZ", [6, 10], [], False), ("RX", [5], [0.0], False),
           ("RY", [0], [0.37], False), ("PauliY", [7], [], False), ("RZ", [7], [0.0], False)]
    # RX on wire 1 and wire 5 share parameter 0; RY(0.37) on wire 0 stays fixed
    pidx = [-1, 0, 1, 2, 3, -1, 4, 5, 6, 7, -1, -1, 8, -1, 0, -1, -1, 9]
    obs = [("X", 1), ("Y", 4), ("Z", 10)]
    t = N.Template(n, ops, pidx, obs)
    assert t.num_params == 10
    thetas = [rng.uniform(0, 2 * np.pi, 10) for _ in range(2)]
    ev, jac = N.batch_run([t, t], thetas)
    for c in range(2):
        e, g = oracle_run(n, ops, pidx, thetas[c], obs)
        close128(ev[c], e)
        close128(jac[c], g)


def test_batch_matches_serial_and_mixes_templates():
    """Uneven circuits of several templates in one launch equal the
    one-circuit-at-a-time form, and each circuit's result does not depend on
    what else is in the batch."""
    rng = np.random.default_rng(5)
    sizes = [8, 14, 10, 12, 9, 14, 11, 13]
    tm = {n: N.Template(n, *ring_template(n), [("Z", 0)]) for n in set(sizes)}
    ts = [tm[n] for n in sizes * 3]
    thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for t in ts]
    ev, jac = N.batch_run(ts, thetas)
    ev_s, jac_s = N.batch_run(ts, thetas, serial=True)
    close128(ev, ev_s)
    for a, b in zip(jac, jac_s):
        close128(a, b)
    # alone: the first circuit by itself
    ev1, jac1 = N.batch_run(ts[:1], thetas[:1])
    close128(ev1[0], ev[0])
    close128(jac1[0], jac[0])


def test_batch_gradient_central_differences():
    rng = np.random.default_rng(3)
    n = 10
    ops, pidx = ring_template(n, layers=2)
    t = N.Template(n, ops, pidx, [("Z", 0)])
    theta = rng.uniform(0, 2 * np.pi, t.num_params)
    _, jac = N.batch_run([t], [theta])
    h = 1e-4
    pick = rng.choice(t.num_params, size=8, replace=False)
    plus, minus = [], []
    for j in pick:
        tp, tm_ = theta.copy(), theta.copy()
        tp[j] += h
        tm_[j] -= h
        plus.append(tp)
        minus.append(tm_)
    ep, _ = N.batch_run([t] * len(pick), plus, grad=False)
    em, _ = N.batch_run([t] * len(pick), minus, grad=False)
    fd = (ep - em) / (2 * h)
    assert np.max(np.abs(fd - jac[0][pick])) < 1e-7


def test_batch_without_gradient_and_fixed_only():
    n = 10
    ops = [("Hadamard", [w], [], False) for w in range(n)] + [("CZ", [0, 1], [], False)]
    # <X5 X7> = 1 exactly (wires outside the CZ stay |+>); <X0> = 0 exactly
    t = N.Template(n, ops, [-1] * len(ops), [("X", 5), ("X", 7)])
    assert t.num_params == 0
    ev, jac = N.batch_run([t, t], [np.zeros(0), np.zeros(0)])
    e, _ = oracle_run(n, ops, [-1] * len(ops), np.zeros(0), [("X", 5), ("X", 7)])
    close128(ev, [e, e])
    t0 = N.Template(n, ops, [-1] * len(ops), [("X", 0)])
    ev0, _ = N.batch_run([t0], [np.zeros(0)], grad=False)
    assert abs(ev0[0]) < 1e-14
    assert jac == [] or all(j.size == 0 for j in jac)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batch_plan_graph_replays_match_eager(dtype):
    """A batch plan (captured HIP graph, autotuned grid) replayed several
    times with new parameters, with eager batched launches and other device
    work in between: every replay equals the eager lgh_batch_run on the same
    parameters bit for bit (same kernels, one workgroup per circuit)."""
    rng = np.random.default_rng(808)
    widths = [9, 12, 10, 14, 9, 16, 11, 13]
    tm = {w: N.Template(w, *ring_template(w, layers=2), [("Z", 0)]) for w in set(widths)}
    ts = [tm[w] for w in widths]
    plan = N.BatchPlan(ts, grad=True, dtype=dtype)
    info = plan.info()
    assert info["grid_per_cu"] in (1, 2) and min(info["tuned_ms"]) > 0
    other = N.State(14, dtype).reset()
    for rep in range(4):
        thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for t in ts]
        ev_p, jac_p = plan.run(thetas)
        ev_e, jac_e = N.batch_run(ts, thetas, dtype=dtype)
        np.testing.assert_array_equal(ev_p, ev_e)
        for a, b in zip(jac_p, jac_e):
            np.testing.assert_array_equal(a, b)
        other.apply([("RX", [rep % 14], [0.3], False), ("CNOT", [0, 5], [], False)])
    # the first circuit against the oracle
    e, g = oracle_run(widths[0], *ring_template(widths[0], layers=2), thetas[0], [("Z", 0)])
    (close128 if dtype == np.complex128 else close64)(ev_p[0], e)
    (close128 if dtype == np.complex128 else close64)(jac_p[0], g)
    plan.close()


def test_batch_plan_without_gradient():
    rng = np.random.default_rng(809)
    t = N.Template(10, *ring_template(10), [("X", 3)])
    plan = N.BatchPlan([t] * 5, grad=False)
    thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for _ in range(5)]
    ev, jac = plan.run(thetas)
    assert jac is None
    ev_e, _ = N.batch_run([t] * 5, thetas, grad=False)
    np.testing.assert_array_equal(ev, ev_e)
    plan.close()


def test_batch_15_to_20_qubits_match_oracle():
    """Circuits at the top of the batched workload's width range (15-20
    qubits) in one batch, complex128, against the oracle."""
    rng = np.random.default_rng(1520)
    widths = [20, 15, 17, 19]
    tm = {w: N.Template(w, *ring_template(w), [("Z", 0)]) for w in widths}
    ts = [tm[w] for w in widths]
    thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for t in ts]
    ev, jac = N.batch_run(ts, thetas)
    for c, w in enumerate(widths):
        e, g = oracle_run(w, *ring_template(w), thetas[c], [("Z", 0)])
        close128(ev[c], e)
        close128(jac[c], g)


def test_batch_c64_multi_circuit_matches_oracle():
    """A complex64 batch of several circuits of different widths against the
    complex128 oracle (within 1e-5)."""
    rng = np.random.default_rng(6464)
    widths = [9, 14, 11, 12, 10, 13, 16]
    tm = {w: N.Template(w, *ring_template(w), [("Z", 0)]) for w in set(widths)}
    ts = [tm[w] for w in widths]
    thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for t in ts]
    ev, jac = N.batch_run(ts, thetas, dtype=np.complex64)
    for c, w in enumerate(widths):
        e, g = oracle_run(w, *ring_template(w), thetas[c], [("Z", 0)])
        close64(ev[c], e)
        close64(jac[c], g)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batch_lean_kernel_matches_general_and_oracle(dtype):
    """Batches whose templates hold only lean gate forms (the RX/RY/RZ +
    CNOT ring) run in the lean batched kernel; a template with controlled
    rotations (CRY) sends the whole batch to the general kernel.  Both, with
    the lean kernels on and off, match the oracle."""
    close = close128 if dtype == np.complex128 else close64
    rng = np.random.default_rng(77)
    ring = {n: N.Template(n, *ring_template(n), [("Z", 0)]) for n in (9, 12)}
    ops = [("RX", [0], [0.0], False), ("CRY", [0, 5], [0.0], False), ("CNOT", [5, 2], [], False),
           ("RZ", [2], [0.0], False), ("RY", [7], [0.0], False)]
    mixed = N.Template(10, ops, [0, 1, -1, 2, 3], [("Z", 2)])
    for batch in ([ring[9], ring[12], ring[9]], [ring[12], mixed, ring[9]]):
        thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for t in batch]
        got = {}
        try:
            for lean in (True, False):
                N.set_lean_passes(lean)
                plan = N.BatchPlan(batch, grad=True, dtype=dtype)
                got[lean] = plan.run(thetas)
                plan.close()
        finally:
            N.set_lean_passes(True)
        for lean in (True, False):
            ev, jac = got[lean]
            for c, t in enumerate(batch):
                n = t.num_qubits
                if t is mixed:
                    e, g = oracle_run(n, ops, [0, 1, -1, 2, 3], thetas[c], [("Z", 2)])
                else:
                    e, g = oracle_run(n, *ring_template(n), thetas[c], [("Z", 0)])
                close(ev[c], e)
                close(jac[c], g)
