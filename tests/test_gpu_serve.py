"""Serving engine (lgh_server_*): results against the oracle, the same result
alone and in a mix of requests, memory admission."""
import numpy as np
import pytest

from _tol import close128
from test_gpu_batch import oracle_run

pytestmark = pytest.mark.gpu

N = pytest.importorskip("pennylane_lightning_gpu._native")


def brick(n, layers, rng):
    """Per layer RY, RZ on every wire, then CZ on even and odd pairs; every
    rotation its own parameter."""
    ops, pidx, k = [], [], 0
    for _ in range(layers):
        for w in range(n):
            for g in ("RY", "RZ"):
                ops.append((g, [w], [float(rng.uniform(0, 2 * np.pi))], False))
                pidx.append(k)
                k += 1
        for s in (0, 1):
            for w in range(s, n - 1, 2):
                ops.append(("CZ", [w, w + 1], [], False))
                pidx.append(-1)
    return ops, pidx


def test_serve_results_match_oracle():
    rng = np.random.default_rng(1)
    srv = N.Server(num_streams=3)
    cases = []
    for n, layers, grad in [(6, 2, True), (10, 3, False), (14, 2, True), (21, 2, True), (12, 4, True)]:
        ops, pidx = brick(n, layers, rng)
        obs = [("Z", 0)] if n % 2 == 0 else [("Z", 0), ("X", n - 1)]
        cases.append((n, ops, pidx, obs, grad, srv.submit(n, ops, obs, grad)))
    for n, ops, pidx, obs, grad, t in cases:
        e, g, tm = srv.wait(t)
        theta = np.array([o[2][0] for o, p in zip(ops, pidx) if p >= 0])
        e_ref, g_ref = oracle_run(n, ops, pidx, theta, obs)
        close128(e, e_ref)
        if grad:
            close128(g, g_ref)
        else:
            assert g is None
        assert tm[0] >= 0 and tm[1] >= 0 and tm[2] >= tm[1]
    st = srv.stats()
    assert st["batched_requests"] + st["single_requests"] == len(cases)
    srv.close()


def test_serve_same_result_alone_and_in_mix():
    rng = np.random.default_rng(2)
    probe_small = brick(12, 3, rng)[0]
    probe_large = brick(22, 2, rng)[0]
    srv = N.Server(num_streams=4)
    alone = []
    for n, ops in ((12, probe_small), (22, probe_large)):
        alone.append(srv.wait(srv.submit(n, ops, [("Z", 0)], True)))
    tickets, probes = [], {}
    for i in range(40):
        n = int(rng.choice([8, 10, 12, 16, 21, 22]))
        ops = brick(n, int(rng.integers(1, 4)), rng)[0]
        tickets.append(srv.submit(n, ops, [("Z", 0)], bool(i % 2)))
        if i == 17:
            probes["small"] = srv.submit(12, probe_small, [("Z", 0)], True)
        if i == 23:
            probes["large"] = srv.submit(22, probe_large, [("Z", 0)], True)
    mixed = {k: srv.wait(t) for k, t in probes.items()}
    for t in tickets:
        srv.wait(t)
    for (e, g, _), key in zip(alone, ("small", "large")):
        assert e == mixed[key][0]
        assert np.array_equal(g, mixed[key][1])
    srv.close()


def test_serve_memory_admission():
    budget = 3 << 26  # 192 MiB
    srv = N.Server(memory_budget=budget, num_streams=2, batch_max_qubits=16)
    rng = np.random.default_rng(3)
    big = brick(23, 1, rng)[0]  # one 23-qubit state is 128 MiB; with the gradient 384 MiB
    with pytest.raises(N.LghError, match="budget"):
        srv.submit(23, big, [("Z", 0)], True)
    # 128 MiB each without the gradient: admitted one at a time
    ts = [srv.submit(23, big, [("Z", 0)], False) for _ in range(3)]
    res = [srv.wait(t) for t in ts]
    assert res[0][0] == res[1][0] == res[2][0]
    assert srv.stats()["budget"] == budget
    srv.close()


def test_serve_failure_injection_leaves_other_requests_intact():
    """A piece of work that fails mid-stream (forced from the host before it
    launches) reports its error to its own requests only; every other request
    queued around it completes with the results it gets alone."""
    rng = np.random.default_rng(4)
    srv = N.Server(num_streams=2, batch_max_qubits=12)
    large = brick(16, 2, rng)[0]
    small = brick(10, 2, rng)[0]
    ref_large = srv.wait(srv.submit(16, large, [("Z", 0)], True))
    ref_small = srv.wait(srv.submit(10, small, [("Z", 0)], True))
    srv.inject_failures(1)
    bad = srv.submit(16, large, [("Z", 0)], True)  # the next piece of work: fails
    with pytest.raises(N.LghError, match="injected failure"):
        srv.wait(bad)
    tickets = [(srv.submit(16, large, [("Z", 0)], True), ref_large) for _ in range(3)]
    tickets += [(srv.submit(10, small, [("Z", 0)], True), ref_small) for _ in range(5)]
    for t, ref in tickets:
        e, g, _ = srv.wait(t)
        assert e == ref[0]
        assert np.array_equal(g, ref[1])
    srv.close()


def test_serve_concurrent_batches_match_single_batch():
    """Small circuits served by several batches at once give the results of
    one batch at a time, bit for bit."""
    rng = np.random.default_rng(5)
    reqs = [(n, brick(n, 2, rng)[0]) for n in rng.choice([9, 10, 11, 12], size=60)]
    out = []
    for limit in (1, 3):
        srv = N.Server(num_streams=4, batch_max_qubits=12)
        srv.set_limits(limit, 8)  # at most 8 circuits per batch: many batches
        ts = [srv.submit(int(n), ops, [("Z", 0)], True) for n, ops in reqs]
        out.append([srv.wait(t) for t in ts])
        assert srv.stats()["batches"] >= len(reqs) // 8
        srv.close()
    for (e1, g1, _), (e2, g2, _) in zip(*out):
        assert e1 == e2
        assert np.array_equal(g1, g2)
