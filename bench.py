#!/usr/bin/env python
"""Benchmark of the lightning.gpu hot path on one MI355X.

Workload (metric "28q adjoint-Jacobian s"): 28 qubits, complex128, 6 layers of
RY/RZ on every wire plus a CZ brick (498 ops, 336 parameters, seeded angles),
observable <Z_0 Z_27>.  One step = re-execute the forward circuit from |0..0>
and compute the 1 x 336 adjoint Jacobian through the pybind11 API, as the
device's adjoint_jacobian does.  Inputs are resident in HBM; the reported
value is seconds per step (lower is better).

Also reported: the 30-qubit single-gate sweep (RX on each target 0..29,
GB/s = 2 * state bytes / kernel time), the roofline of the dominant kernel
(the fused adjoint step), and the CPU oracle timed on a bounded sample of the
same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "pennylane-lightning-gpu_amd")
for _p in (ROOT, PKG_DIR):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def kernel_sources_sha16():
    """Hash of the sources of the fused-pass kernels (device code and the
    headers it includes): a PMC traffic figure measured on other sources is
    stale."""
    import hashlib

    h = hashlib.sha256()
    for rel in ("lgh_fused.hip", "lgh_fused_dev.hpp", "lgh_internal.hpp"):
        with open(os.path.join(PKG_DIR, "csrc", "lgh", rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--layers", type=int, default=6)
    ap.add_argument("--sweep-qubits", type=int, default=30)
    ap.add_argument("--sweep-reps", type=int, default=10)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1337)
    ap.add_argument("--tile-bits", type=int, nargs=2, default=None, metavar=("FWD", "ADJ"),
                    help="register bits of the fused forward / adjoint passes (library default: 4, and 4 "
                         "for the adjoint from 24 qubits, else 3; adjoint 0 keeps that default)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tile heights and grids instead of autotuning them")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-measure", action="store_true")
    ap.add_argument("--serve-arrivals", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--serve-batch-max", type=int, default=16,
                    help="circuits of at most this many qubits are batched; larger ones run on their own states")
    ap.add_argument("--serve-plain-sample", type=int, default=10,
                    help="requests per (qubits, gradient) class timed in the plain per-circuit form")
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=None, metavar="RATE",
                    help="offered rates (requests/s) of a sustainable-rate sweep")
    ap.add_argument("--serve-sweep-arrivals", type=int, default=600)
    ap.add_argument("--serve-only", action="store_true", help="run only the serving workload")
    ap.add_argument("--published-wires", type=int, nargs="*", default=[21, 23, 25, 26, 27],
                    help="wire counts of the reference's published StronglyEntanglingLayers workload")
    ap.add_argument("--no-published", action="store_true")
    ap.add_argument("--published-only", action="store_true", help="run only the published workload")
    ap.add_argument("--batch-circuits", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-serial-sample", type=int, default=0,
                    help="circuits timed one after another for the per-circuit comparison (0: all of them)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run device-to-device copies of 1 GiB (PMC byte-counter calibration)")
    return ap.parse_args()


def copy_calibration():
    """1 GiB device copies: the known-size reference for the FETCH_SIZE /
    WRITE_SIZE counters (read 1 GiB + write 1 GiB per copy)."""
    from pennylane_lightning_gpu import _native as N

    a, b = N.State(26).reset(), N.State(26)  # 2^26 * 16 B = 1 GiB each
    for _ in range(3):
        b.copy_from(a)
    # the library's own streaming kernel of known size: y += a x reads 2 GiB, writes 1 GiB
    for _ in range(3):
        b.axpy(0.5, a)
    b.synchronize()
    print(json.dumps({"copy_calibration": {"copy": {"bytes_read": 1 << 30, "bytes_written": 1 << 30},
                                           "k_axpy": {"bytes_read": 2 << 30, "bytes_written": 1 << 30},
                                           "repeats": 3}}))


class StreamTimer:
    """HIP events recorded on the library stream (through torch's external
    stream wrapper, which shares the process's HIP runtime)."""

    def __init__(self, stream_ptr):
        import torch

        self.torch = torch
        self.stream = torch.cuda.ExternalStream(stream_ptr) if stream_ptr else torch.cuda.default_stream()

    def time(self, fn, reps=1):
        t = self.torch
        start = t.cuda.Event(enable_timing=True)
        end = t.cuda.Event(enable_timing=True)
        start.record(self.stream)
        for _ in range(reps):
            fn()
        end.record(self.stream)
        end.synchronize()
        return start.elapsed_time(end) / reps  # ms


def bench_batched(N, args, dtype):
    """Batched workload: circuits with n ~ U{8..20} qubits (seeded), each 3 x
    [RX, RY, RZ on every wire; CNOT ring] (12n gates, 9n parameters), <Z_0>
    and its gradient.  One step = all circuits in one persistent launch,
    replayed from a batch plan (captured HIP graph, grid autotuned at plan
    creation).  Compared with the per-circuit launch form on a sample of the
    same circuits.  Roofline of the batched kernel: algorithmic bytes
    (2 F + 4 R + 4) S_c per circuit (F forward and R reverse passes over
    states of S_c bytes, the observable pass and the initial state) over the
    kernel time; states of up to 2^20 amplitudes may stay in L2 / MALL
    between passes, so HBM moves fewer bytes than this."""
    import numpy as np

    from pennylane_lightning_gpu.workloads import batch_widths, ring_template

    widths = batch_widths(args.batch_circuits, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    tm = {}
    for w in sorted(set(widths)):
        ops, pidx = ring_template(w)
        tm[w] = N.Template(w, ops, pidx, [("Z", 0)])
    ts = [tm[w] for w in widths]
    thetas = [rng.uniform(0, 2 * np.pi, t.num_params) for t in ts]
    elem = 16 if np.dtype(dtype) == np.complex128 else 8
    alg_bytes = 0
    for t in ts:
        f, r, q = t.passes()
        alg_bytes += (2 * f + 4 * r + 4) * elem * (1 << q)
    name = np.dtype(dtype).name
    log(f"[bench] batched {name}: {len(ts)} circuits, n in [{min(widths)}, {max(widths)}], forward + gradient")
    t0 = time.perf_counter()
    plan = N.BatchPlan(ts, grad=True, dtype=dtype)
    t_plan = time.perf_counter() - t0
    plan.run(thetas)  # warm-up
    N.profile_enable(True)
    steps = []
    for i in range(args.batch_steps):
        t0 = time.perf_counter()
        ev, jac = plan.run(thetas)
        steps.append(time.perf_counter() - t0)
        log(f"[bench] batched {name} step {i}: {steps[-1] * 1e3:.1f} ms")
    kn, kms = N.profile_read("batch")
    N.profile_enable(False)
    k = len(ts) if args.batch_serial_sample <= 0 else min(args.batch_serial_sample, len(ts))
    t0 = time.perf_counter()
    ev_s, jac_s = N.batch_run(ts[:k], thetas[:k], serial=True, dtype=dtype)
    t_serial = time.perf_counter() - t0
    diff = max(float(np.max(np.abs(ev[:k] - ev_s))),
               max(float(np.max(np.abs(a - b))) for a, b in zip(jac[:k], jac_s)))
    st = np.array(steps)
    cps = len(ts) / float(np.median(st))
    # the sample's share of the full batch's work, by circuit cost ~ n 2^n
    cost = np.array([w * 2.0 ** w for w in widths])
    serial_full = t_serial * cost.sum() / cost[:k].sum()
    kernel_ms = kms / max(1, kn)
    out = {
        "circuits": len(ts), "qubits": [min(widths), max(widths)], "dtype": name,
        "circuits_per_s": round(cps, 1),
        "step_ms_p50": round(float(np.percentile(st, 50)) * 1e3, 2),
        "step_ms_p99": round(float(np.percentile(st, 99)) * 1e3, 2),
        "kernel_ms_per_step": round(kernel_ms, 2),
        "plan": plan.info() | {"create_s": round(t_plan, 3)},
        "roofline": {"bound": "hbm", "kernel": f"k_batch<{'double' if elem == 16 else 'float'}>",
                     "algorithmic_bytes": int(alg_bytes),
                     "achieved": round(alg_bytes / (kernel_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)},
        "per_circuit_form": {"sample": k, "seconds": round(t_serial, 3),
                             "circuits_per_s_est": round(len(ts) / serial_full, 1),
                             "max_abs_diff_vs_batched": diff},
        "speedup_vs_per_circuit": round(serial_full / float(np.median(st)), 1),
    }
    plan.close()
    for t in tm.values():
        t.close()
    return out


def bench_measurements(N, args):
    """Measurement workloads, complex128: the state of 8 x [RX, RY, RZ on every
    wire; CNOT ladder] from |0..0>.  On 24 qubits: <H> of 256 random Pauli
    words (1-4 factors); a dense 2-wire expectation <Z0 Z23>; the variance of
    a random 3-wire Hermitian (two dense expectations, as the reference
    computes it); marginal probabilities of 10 wires; 10^6 seeded samples.  On
    20 qubits: <H> of the open-chain TFIM (J = 1, h = 0.7) as a resident CSR.
    Times are HIP events on the state's stream, median of repeats.  Last, the
    torch.autograd form: forward and backward of 256 circuits of the 12-qubit
    ring template (host wall clock)."""
    import numpy as np

    from pennylane_lightning_gpu.workloads import ladder, random_pauli_hamiltonian, tfim_csr

    out = {}
    n = 24
    S = 16 << n
    psi = N.State(n).reset()
    psi.apply(ladder(n, 8, seed=args.seed))
    timer = StreamTimer(psi.stream())
    words, wires, coeffs = random_pauli_hamiltonian(n, 256, seed=args.seed)
    psi.expval_pauli_words(words, wires, coeffs)  # warm
    ms = sorted(timer.time(lambda: psi.expval_pauli_words(words, wires, coeffs)) for _ in range(5))[2]
    # one read of the state per launch: tile passes of up to 32 words, pair
    # passes of up to 8 words sharing a flip mask
    plan = N.pauli_plan(n, words, wires)
    passes = plan["tile_launches"] + plan["pair_launches"]
    out["pauli_words"] = {"qubits": n, "words": len(words), "ms": round(ms, 3), "plan": plan,
                          "state_reads": passes, "bytes": passes * S,
                          "GBps": round(passes * S / (ms * 1e-3) / 1e9, 1),
                          "words_per_s": round(len(words) / (ms * 1e-3), 1)}
    # dense expectation <Z0 Z23> and a 3-wire Hermitian (one read pass each), and
    # the variance of the latter as the reference computes it (two passes)
    Z = np.diag([1.0, -1.0])
    zz = np.kron(Z, Z).astype(np.complex128)
    psi.expval_matrix([0, n - 1], zz)  # warm
    ms = sorted(timer.time(lambda: psi.expval_matrix([0, n - 1], zz)) for _ in range(5))[2]
    out["expval_dense"] = {"qubits": n, "wires": 2, "ms": round(ms, 3), "bytes": S,
                           "GBps": round(S / (ms * 1e-3) / 1e9, 1)}
    rng = np.random.default_rng(args.seed)
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    Hm = (A + A.conj().T) / 2
    Hm2 = Hm @ Hm
    w3 = [2, 11, 19]
    psi.expval_matrix(w3, Hm)  # warm
    ms_var = sorted(timer.time(lambda: (psi.expval_matrix(w3, Hm), psi.expval_matrix(w3, Hm2)))
                    for _ in range(5))[2]
    out["variance_hermitian3"] = {"qubits": n, "wires": 3, "ms": round(ms_var, 3), "bytes": 2 * S,
                                  "GBps": round(2 * S / (ms_var * 1e-3) / 1e9, 1)}
    pw = list(range(0, n, 2))[:10]
    psi.probs(pw)  # warm
    ms = sorted(timer.time(lambda: psi.probs(pw)) for _ in range(5))[2]
    out["probabilities"] = {"qubits": n, "wires": len(pw), "ms": round(ms, 3), "bytes": S,
                            "GBps": round(S / (ms * 1e-3) / 1e9, 1)}
    shots = 10 ** 6
    psi.sample(shots, args.seed)  # warm
    ms = sorted(timer.time(lambda: psi.sample(shots, args.seed)) for _ in range(3))[1]
    sbytes = S + 2 * 8 * (1 << n) + 8 * shots
    out["sampling"] = {"qubits": n, "shots": shots, "ms": round(ms, 3), "bytes": sbytes,
                       "GBps": round(sbytes / (ms * 1e-3) / 1e9, 1)}
    del psi
    n = 20
    st = N.State(n).reset()
    st.apply(ladder(n, 8, seed=args.seed))
    indptr, indices, data = tfim_csr(n)
    csr = N.CSR(indptr, indices, data)
    csr.expval(st)  # warm
    timer = StreamTimer(st.stream())
    ms = sorted(timer.time(lambda: csr.expval(st)) for _ in range(5))[2]
    nnz, NN = len(data), 1 << n
    cbytes = 24 * nnz + 8 * (NN + 1) + 16 * NN
    out["csr_tfim"] = {"qubits": n, "nnz": nnz, "ms": round(ms, 3), "bytes": cbytes,
                       "GBps": round(cbytes / (ms * 1e-3) / 1e9, 1)}
    del st, csr
    # torch.autograd over a batch: 256 circuits of the 12-qubit ring template,
    # forward (one batched launch with the Jacobians) and backward
    import time

    import torch

    from pennylane_lightning_gpu.torch_layer import batched_expval
    from pennylane_lightning_gpu.workloads import ring_template

    ops, pidx = ring_template(12)
    t = N.Template(12, ops, pidx, [("Z", 0)])
    p = torch.tensor(np.random.default_rng(args.seed).uniform(0, 2 * np.pi, (256, t.num_params)),
                     dtype=torch.float64, requires_grad=True)
    batched_expval(p, t).sum().backward()  # warm
    times = []
    for _ in range(5):
        p.grad = None
        t0 = time.perf_counter()
        batched_expval(p, t).sum().backward()
        times.append(time.perf_counter() - t0)
    ms = sorted(times)[2] * 1e3
    out["torch_batched"] = {"qubits": 12, "circuits": 256, "params": t.num_params, "ms_fwd_bwd": round(ms, 3),
                            "circuits_per_s": round(256 / (ms * 1e-3), 1)}
    t.close()
    return out


# The reference's published adjoint benchmark: qml.jacobian wall time of
# StronglyEntanglingLayers with <Z_w> on every wire, complex128, one A100,
# read off the log-scale bar chart doc/_static/lightning_gpu_initial_bm.png
# (.github/CHANGELOG.md:603-640) to about +-5 %.  Keys: (wires, layers).
PUBLISHED_A100_S = {(21, 1): 2.7, (21, 2): 4.5, (23, 1): 5.6, (23, 2): 8.0,
                    (25, 1): 16.0, (25, 2): 19.0, (27, 1): 57.0, (27, 2): 72.0}


def bench_published(N, ops_mod, args):
    """The reference's published workload shape on this build, through the
    pybind11 API as the device drives it: reset, the StronglyEntanglingLayers
    forward, <Z_w> for every wire, and the n x 3nL adjoint Jacobian of all n
    observables.  One warm-up, median of 3.  The fused adjoint passes of the
    timed runs are timed with HIP events: each observable's sweep forms its
    observable state inside its first pass (3 state bytes) and moves 4 state
    bytes in every other pass."""
    import numpy as np

    from pennylane_lightning_gpu.workloads import strongly_entangling

    z = np.diag([1.0, -1.0]).astype(np.complex128).ravel()
    rows = []
    for n in args.published_wires:
        for layers in (1, 2):
            ops = strongly_entangling(n, layers, seed=args.seed)
            n_par = 3 * n * layers
            S = 16 << n
            sv = ops_mod.LightningGPU_C128(n)
            adj = ops_mod.AdjointJacobianGPU_C128()
            obs = [ops_mod.NamedObsGPU_C128("PauliZ", [w]) for w in range(n)]
            opsl = adj.create_ops_list([o[0] for o in ops], [np.array(o[2], dtype=np.float64) for o in ops],
                                       [o[1] for o in ops], [o[3] for o in ops],
                                       [np.array([], complex) for _ in ops])
            names, wires, invs, params = ([o[0] for o in ops], [o[1] for o in ops], [o[3] for o in ops],
                                          [o[2] for o in ops])

            def run():
                sv.resetGPU()
                sv.apply(names, wires, invs, params)
                ev = [sv.ExpectationValue("PauliZ", [w], [], z) for w in range(n)]
                return ev, np.asarray(adj.adjoint_jacobian(sv, obs, opsl, list(range(n_par))))

            run()
            N.check(N.lib().lgh_device_synchronize(0))
            N.profile_enable(True)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                ev, jac = run()
                N.check(N.lib().lgh_device_synchronize(0))
                ts.append(time.perf_counter() - t0)
            a_n, a_ms = N.profile_read("fused_adjoint")
            f_n, f_ms = N.profile_read("fused_pass")
            N.profile_enable(False)
            sec = sorted(ts)[1]
            a_bytes = (4 * a_n - 3 * n) * S / max(1, a_n)  # per launch, averaged over 3 runs
            pub = PUBLISHED_A100_S.get((n, layers))
            rows.append({
                "wires": n, "layers": layers, "params": n_par, "observables": n, "s": round(sec, 4),
                "published_a100_s": pub, "ratio_vs_published": round(pub / sec, 1) if pub else None,
                "adjoint_launches": a_n // 3, "adjoint_ms_per_launch": round(a_ms / max(1, a_n), 4),
                "adjoint_bytes_per_launch": int(a_bytes),
                "adjoint_GBps": round(a_bytes / (a_ms / max(1, a_n) * 1e-3) / 1e9, 1) if a_n else None,
                "adjoint_ms": round(a_ms / 3, 2), "forward_launches": f_n // 3, "forward_ms": round(f_ms / 3, 2),
                "jacobian_checksum": float(jac.sum()), "expval_checksum": float(np.sum(ev)),
            })
            log(f"[bench] published shape {n} wires x {layers} layers: {sec:.3f} s "
                f"(A100 published {pub} s), {a_n // 3} adjoint passes")
            del sv, adj, obs, opsl
            N.check(N.lib().lgh_pool_trim(0))
    return {"workload": "StronglyEntanglingLayers (Rot expanded to RZ RY RZ, CNOT ranges (l mod (n-1)) + 1), "
                        "<Z_w> on every wire, complex128, forward + expectation values + adjoint Jacobian",
            "published_note": "A100 figures are qml.jacobian wall times including PennyLane overhead, read off "
                              "the reference's bar chart (about +-5 %); this build's times exclude PennyLane",
            "rows": rows}


def serving_requests(args, arrivals, rate, seed, max_qubits=None):
    """Config-5 arrivals: Poisson at ``rate`` per second (seeded exponential
    gaps); n from {10, 14, 18, 22, 26, 30} with weights 8:6:4:3:2:1 (classes
    above ``max_qubits`` dropped); layers ~ U{2..8} of [RY, RZ on every wire;
    CZ on even then odd pairs]; half the requests ask for <Z_0>, half for its
    gradient."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = np.array([10, 14, 18, 22, 26, 30])
    weights = np.array([8, 6, 4, 3, 2, 1], dtype=float)
    if max_qubits is not None:
        weights = np.where(sizes <= max_qubits, weights, 0.0)
    reqs = []
    t = 0.0
    for i in range(arrivals):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice(sizes, p=weights / weights.sum()))
        layers = int(rng.integers(2, 9))
        ops = []
        for _ in range(layers):
            for w in range(n):
                ops.append(("RY", [w], [float(rng.uniform(0, 2 * np.pi))], False))
                ops.append(("RZ", [w], [float(rng.uniform(0, 2 * np.pi))], False))
            for s0 in (0, 1):
                for w in range(s0, n - 1, 2):
                    ops.append(("CZ", [w, w + 1], [], False))
        reqs.append((t, n, ops, bool(i % 2)))
    return reqs


def serve_stream(N, reqs, streams, batch_max, batch_wait_us=None):
    """Submits ``reqs`` on their schedule to a fresh server and drains it.
    Returns throughput and latency figures (queueing counted from the
    scheduled arrival)."""
    import numpy as np

    srv = N.Server(num_streams=streams, batch_max_qubits=batch_max)
    if batch_wait_us is not None:
        srv.set_batching(64, batch_wait_us)
    # warm the template cache and the allocator with one request per class
    for n in sorted({r[1] for r in reqs}):
        srv.wait(srv.submit(int(n), [("RY", [0], [0.1], False)], [("Z", 0)], True))
    # the client packs every request before the clock starts (op arrays are
    # the client's work, not the engine's)
    packed = [srv.prepare(n, ops, [("Z", 0)], grad) for _, n, ops, grad in reqs]
    tickets = [None] * len(reqs)
    late = np.zeros(len(reqs))  # submission after the scheduled arrival (host side), ms
    t0 = time.perf_counter()
    for i, (ta, n, ops, grad) in enumerate(reqs):
        if i and i % 5000 == 0:
            log(f"[bench] serving: {i} requests submitted at {time.perf_counter() - t0:.1f} s")
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        late[i] = max(0.0, (time.perf_counter() - t0) - ta) * 1e3
        tickets[i] = srv.submit_prepared(packed[i])
    t_sub = time.perf_counter() - t0
    lat = {}
    for i, ((ta, n, ops, grad), tk) in enumerate(zip(reqs, tickets)):
        _, _, tm = srv.wait(tk)
        lat.setdefault(int(n), []).append((tm[0] + late[i], tm[1], tm[2] + late[i]))
        if (i + 1) % 2000 == 0:
            log(f"[bench] serving: {i + 1} of {len(reqs)} requests done at {time.perf_counter() - t0:.1f} s")
    t_all = time.perf_counter() - t0
    st = srv.stats()
    srv.close()

    def pct(v, q):
        return round(float(np.percentile(v, q)), 2)

    per = {}
    alltm = []
    for n in sorted(lat):
        a = np.array(lat[n])
        alltm.append(a)
        per[str(n)] = {"count": len(a), "latency_ms_p50": pct(a[:, 2], 50), "latency_ms_p99": pct(a[:, 2], 99),
                       "queue_ms_p50": pct(a[:, 0], 50), "queue_ms_p99": pct(a[:, 0], 99),
                       "run_ms_p50": pct(a[:, 1], 50), "run_ms_p99": pct(a[:, 1], 99)}
    a = np.concatenate(alltm)
    offered = len(reqs) / reqs[-1][0] if reqs[-1][0] > 0 else float("inf")
    return {
        "arrivals": len(reqs), "offered_rate": round(offered, 1), "streams": streams, "batch_max_qubits": batch_max,
        "submit_window_s": round(t_sub, 3), "drain_s": round(t_all, 3),
        "submit_lateness_ms_p99": pct(late, 99),
        "achieved_circuits_per_s": round(len(reqs) / t_all, 1),
        "latency_ms_p50": pct(a[:, 2], 50), "latency_ms_p99": pct(a[:, 2], 99),
        "queue_ms_p50": pct(a[:, 0], 50), "queue_ms_p99": pct(a[:, 0], 99),
        "per_class": per, "engine": st,
    }


def per_circuit_times(N, reqs):
    """The plain per-circuit form: requests one after another, each on its
    own states through the single-circuit entry points (one stream, no
    batching, no overlap).  Mean wall seconds per request by (qubits,
    gradient) class."""
    import numpy as np

    srv = N.Server(num_streams=1, batch_max_qubits=0)
    for n in sorted({r[1] for r in reqs}):
        srv.wait(srv.submit(int(n), [("RY", [0], [0.1], False)], [("Z", 0)], True))
    per = {}
    for _, n, ops, grad in reqs:
        packed = srv.prepare(n, ops, [("Z", 0)], grad)
        t0 = time.perf_counter()
        srv.wait(srv.submit_prepared(packed))
        per.setdefault((n, grad), []).append(time.perf_counter() - t0)
    srv.close()
    return {k: float(np.mean(v)) for k, v in per.items()}


def plain_rate(reqs, times):
    """Circuits per second of the per-circuit form over ``reqs``, from the
    per-class mean times (every class of ``reqs`` must have been timed)."""
    return len(reqs) / sum(times[(n, grad)] for _, n, _, grad in reqs)


def bench_serving(N, args):
    """Serving workload (config 5): ``--serve-arrivals`` Poisson arrivals at
    ``--serve-rate`` per second submitted on schedule and drained; the same
    stream restricted to the classes up to 22 qubits; the plain per-circuit
    form timed on a sample of each stream for the served / per-circuit
    circuits-per-second ratio; with ``--serve-sweep``, a sweep of offered
    rates for the sustainable rate."""
    out = {}
    for label, maxq in (("mix_10_30", None), ("mix_10_22", 22)):
        reqs = serving_requests(args, args.serve_arrivals, args.serve_rate, args.seed, maxq)
        log(f"[bench] serving {label}: {len(reqs)} arrivals at {args.serve_rate}/s over {reqs[-1][0]:.2f} s")
        r = serve_stream(N, reqs, args.serve_streams, args.serve_batch_max)
        # capacity: the same requests all submitted at once
        sat = serve_stream(N, [(0.0,) + q[1:] for q in reqs], args.serve_streams, args.serve_batch_max)
        # per-circuit form: every (qubits, gradient) class timed on a sample
        sample, seen = [], {}
        for q in reqs:
            key = (q[1], q[3])
            if seen.get(key, 0) < args.serve_plain_sample:
                seen[key] = seen.get(key, 0) + 1
                sample.append(q)
        times = per_circuit_times(N, sample)
        plain = plain_rate(reqs, times)
        log(f"[bench] serving {label}: {r['achieved_circuits_per_s']} circuits/s at the offered rate, "
            f"{sat['achieved_circuits_per_s']} at saturation, {plain:.1f} per-circuit form")
        r["saturation"] = {k: sat[k] for k in ("achieved_circuits_per_s", "drain_s", "latency_ms_p50",
                                               "latency_ms_p99", "engine")}
        r["per_circuit_form"] = {"circuits_per_s": round(plain, 1), "sample_per_class": args.serve_plain_sample,
                                 "mean_ms_by_class": {f"{n}q{'_grad' if g else ''}": round(t * 1e3, 3)
                                                      for (n, g), t in sorted(times.items())}}
        r["served_vs_per_circuit"] = round(sat["achieved_circuits_per_s"] / plain, 2)
        out[label] = r
    if args.serve_sweep:
        sweep = []
        for rate in args.serve_sweep:
            reqs = serving_requests(args, args.serve_sweep_arrivals, rate, args.seed + 1)
            log(f"[bench] serving sweep: {len(reqs)} arrivals at {rate}/s")
            r = serve_stream(N, reqs, args.serve_streams, args.serve_batch_max)
            sweep.append({k: r[k] for k in ("offered_rate", "achieved_circuits_per_s", "drain_s", "latency_ms_p50",
                                            "latency_ms_p99", "queue_ms_p50", "queue_ms_p99")}
                         | {"class_10": r["per_class"].get("10")})
        out["rate_sweep"] = sweep
    return out


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(args, ops, n):
    """The CPU oracle (the checker: a C/OpenMP restatement of the reference's
    CPU path) timed on this host, on rank 0 only.

    - The plain 12-qubit configuration, run whole: |0..0>, 192 ops, the 12
      expectation values <Z_w>, the 12 x 144 adjoint Jacobian.
    - The headline workload, summed from measured components: allocating and
      zeroing the state, one forward pass per op class (RY, RZ, CZ), and the
      adjoint sweep's setup and reverse step per op class, each timed at the
      workload's own size and multiplied by that class's op count.  A full
      28-qubit run takes minutes; the component sum takes about 20-40 s.
    """
    import numpy as np

    from oracle import pyoracle as po
    from pennylane_lightning_gpu.workloads import layered_ring

    threads = po.num_threads()
    out = {"unit": "s", "cores": threads, "kind": "port", "cpu_model": cpu_info(),
           "binding": os.environ.get("OMP_PROC_BIND", "unset") + "/" + os.environ.get("OMP_PLACES", "unset")}
    # plain configuration, whole
    c1 = layered_ring(12, 4, seed=args.seed)
    t0 = time.perf_counter()
    psi = po.apply_ops(12, po.zero_state(12), c1)
    zs = [po.expval_matrix(12, psi, [w], np.diag([1.0, -1.0])).real for w in range(12)]
    bras = [po.apply_ops(12, psi, [("PauliZ", [w], [], False)]) for w in range(12)]
    po.adjoint_jacobian(12, psi, bras, c1, list(range(144)))
    out["config1"] = {"workload": "12 qubits c128, 4 x (RX, RY, RZ per wire; CNOT ring), 12 <Z_w>, 12 x 144 adjoint",
                      "cpu_s": round(time.perf_counter() - t0, 4), "checksum": float(np.sum(zs))}
    # headline workload by components
    log(f"[bench] CPU oracle: {n}q components")
    counts = {}
    for o in ops:
        counts[o[0]] = counts.get(o[0], 0) + 1
    th, k = 0.731, 4

    def timed(fn):
        t = time.perf_counter()
        r = fn()
        return time.perf_counter() - t, r

    small = po.zero_state(16)  # loads the library and its thread pool
    po.adjoint_jacobian(16, small, [small], [("RY", [0], [th], False)], [0])
    t_zero, psi = timed(lambda: po.zero_state(n))
    # k gates of each class on wires spread over the register (a gate's cost
    # depends on its target's stride)
    spread = [round(i * (n - 2) / (k - 1)) for i in range(k)]
    gate = {"RY": [("RY", [w], [th], False) for w in spread], "RZ": [("RZ", [w], [th], False) for w in spread],
            "CZ": [("CZ", [w, w + 1], [], False) for w in spread]}
    # forward: apply_ops copies the state, then applies the ops; k ops
    # against none isolates the per-op cost
    t_copy = sorted(timed(lambda: po.apply_ops(n, psi, []))[0] for _ in range(3))[1]
    fwd = {g: max(0.0, (timed(lambda: po.apply_ops(n, psi, seq))[0] - t_copy) / k) for g, seq in gate.items()}
    # adjoint: one trainable RY alone, then followed by k ops of a class (all
    # trainable): the difference is k reverse steps of that class
    bra = po.apply_ops(n, psi, [("PauliZ", [0], [], False), ("PauliZ", [n - 1], [], False)])
    lead = ("RY", [0], [th], False)
    t1 = sorted(timed(lambda: po.adjoint_jacobian(n, psi, [bra], [lead], [0]))[0] for _ in range(2))[0]
    rev = {}
    for g, seq0 in gate.items():
        seq = [lead] + seq0
        tp = list(range(sum(1 for o in seq if o[2])))
        rev[g] = max(0.0, (timed(lambda: po.adjoint_jacobian(n, psi, [bra], seq, tp))[0] - t1) / k)
    setup = max(0.0, t1 - rev["RY"])
    del psi, bra
    total = t_zero + t_copy + setup + sum(counts[g] * (fwd[g] + rev[g]) for g in counts)
    out["value"] = round(total, 2)
    out["components"] = {"zero_state_s": round(t_zero, 3), "forward_copy_s": round(t_copy, 3),
                         "adjoint_setup_s": round(setup, 3),
                         "forward_s_per_op": {g: round(v, 4) for g, v in fwd.items()},
                         "reverse_s_per_op": {g: round(v, 4) for g, v in rev.items()}, "op_counts": counts}
    out["sample"] = (f"{n}q c128 headline summed from components timed at {n} qubits: state zeroing, forward pass "
                     f"and adjoint reverse step per op class (RY, RZ, CZ; four ops each on spread wires), adjoint setup; "
                     f"{len(ops)} ops. Plain 12q configuration timed whole.")
    return out


def gpu_config1(ops_mod, args):
    """The plain 12-qubit configuration on the GPU through the pybind11 API,
    as the device runs it: forward, the 12 <Z_w>, the 12 x 144 adjoint
    Jacobian; median of 5 after a warm-up."""
    import numpy as np

    from pennylane_lightning_gpu.workloads import layered_ring

    c1 = layered_ring(12, 4, seed=args.seed)
    sv = ops_mod.LightningGPU_C128(12)
    adj = ops_mod.AdjointJacobianGPU_C128()
    obs = [ops_mod.NamedObsGPU_C128("PauliZ", [w]) for w in range(12)]
    opsl = adj.create_ops_list([o[0] for o in c1], [np.array(o[2], dtype=np.float64) for o in c1],
                               [o[1] for o in c1], [o[3] for o in c1], [np.array([], complex) for _ in c1])
    z = np.diag([1.0, -1.0]).astype(np.complex128).ravel()

    def run():
        sv.resetGPU()
        sv.apply([o[0] for o in c1], [o[1] for o in c1], [o[3] for o in c1], [o[2] for o in c1])
        [sv.ExpectationValue("PauliZ", [w], [], z) for w in range(12)]
        return adj.adjoint_jacobian(sv, obs, opsl, list(range(144)))

    run()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    return round(sorted(ts)[2], 5)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every box in this setup has one GPU.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    import torch  # noqa: F401  (HIP events)

    from pennylane_lightning_gpu import _native as N
    from pennylane_lightning_gpu import lightning_gpu_qubit_ops as ops_mod
    from pennylane_lightning_gpu.workloads import ry_rz_cz_brick

    if args.tile_bits:
        N.set_tile_bits(*args.tile_bits)
    if args.serve_only:
        print(json.dumps({"serving": bench_serving(N, args)}))
        return
    if args.published_only:
        print(json.dumps({"published_workload": bench_published(N, ops_mod, args)}))
        return
    n, layers = args.qubits, args.layers
    ops = ry_rz_cz_brick(n, layers, seed=args.seed)
    n_par = sum(1 for o in ops if o[2])
    state_bytes = 16 * (1 << n)

    # one-time setup: the autotuner picks tile height and grid of the fused
    # passes for this state size on the headline circuit (outside the timed steps)
    tune = None
    if not args.no_autotune and not args.tile_bits:
        log(f"[bench] autotuning the fused passes on the {n}-qubit headline circuit")
        tune = N.autotune_fused(n, ops)
        log(f"[bench] tuned: forward {tune['forward']}, adjoint {tune['adjoint']}")
    # register bits of the adjoint passes the timed steps use (the library's
    # default: 4 for complex128 states of 24 qubits and more, else 3)
    if tune is not None:
        adj_rb = int(tune["adjoint"]["tile_bits"])
    elif args.tile_bits and args.tile_bits[1]:
        adj_rb = int(args.tile_bits[1])
    else:
        adj_rb = 4 if n >= 24 else 3
    adj_kernel = f"k_fused_adj{adj_rb}<double"  # the lean level follows from the passes

    # ---------------------------------------------------------------- headline: adjoint Jacobian
    sv = ops_mod.LightningGPU_C128(n)
    adj = ops_mod.AdjointJacobianGPU_C128()
    obs = ops_mod.TensorProdObsGPU_C128([ops_mod.NamedObsGPU_C128("PauliZ", [0]),
                                         ops_mod.NamedObsGPU_C128("PauliZ", [n - 1])])
    opsl = adj.create_ops_list([o[0] for o in ops], [np.array(o[2], dtype=np.float64) for o in ops],
                               [o[1] for o in ops], [o[3] for o in ops], [np.array([], complex) for _ in ops])
    names, wires, invs, params = [o[0] for o in ops], [o[1] for o in ops], [o[3] for o in ops], [o[2] for o in ops]
    tp = list(range(n_par))

    def step():
        sv.resetGPU()
        sv.apply(names, wires, invs, params)
        return np.asarray(adj.adjoint_jacobian(sv, [obs], opsl, tp))

    log(f"[bench] {n}q adjoint: {len(ops)} ops, {n_par} params; warmup {args.warmup}")
    for _ in range(args.warmup):
        jac = step()
    N.check(N.lib().lgh_device_synchronize(0))
    # HIP events around every fused-pass launch of the timed steps, recorded
    # by the library on the stream it launches on
    N.profile_enable(True)
    times = []
    for i in range(args.steps):
        N.check(N.lib().lgh_device_synchronize(0))
        t0 = time.perf_counter()
        jac = step()
        N.check(N.lib().lgh_device_synchronize(0))
        times.append(time.perf_counter() - t0)
        log(f"[bench] step {i}: {times[-1]:.3f} s")
    sec = float(np.mean(times))
    jac = np.asarray(jac).ravel()
    adj_n, adj_ms = N.profile_read("fused_adjoint")
    fwd_n, fwd_ms = N.profile_read("fused_pass")
    N.profile_enable(False)

    # ---------------------------------------------------------------- dominant kernel: fused adjoint pass
    # A launch reads and writes lambda and the observable state once: 4 state
    # bytes per launch, whatever the number of gates fused into it.  The
    # observable Z0 Z27 is a Z word, which the first reverse pass of each step
    # forms from lambda instead of reading it: 3 state bytes for that launch.
    adj_avg = adj_ms / max(1, adj_n)
    fused_bytes = (4 * adj_n - args.steps) * state_bytes / max(1, adj_n)  # per launch, averaged
    fused_gbs = fused_bytes / (adj_avg * 1e-3) / 1e9
    fwd_avg = fwd_ms / max(1, fwd_n)
    # HBM bytes per launch from the committed PMC summary, used only if it
    # was measured on the current kernel sources
    traffic, traffic_note, valu = None, None, None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile) and n == 28:
        with open(tfile) as f:
            entry = json.load(f).get(f"k_fused_adj<double>@{n}q", {})
        if entry.get("kernel_sources_sha16") == kernel_sources_sha16() and entry.get("kernels") and \
                all(k.startswith(adj_kernel) for k in entry["kernels"]):
            adj_kernel = " + ".join(entry["kernels"])
            traffic = entry.get("traffic_bytes")
            if entry.get("valu_insts"):
                # VALU issue: a wave64 instruction occupies a SIMD for 4 cycles; 1024 SIMDs
                # at the 2.4 GHz peak engine clock
                issue_s = entry["valu_insts"] * 4 / (1024 * 2.4e9)
                valu = {"insts_per_launch": int(entry["valu_insts"]),
                        "issue_ms_at_peak_clock": round(issue_s * 1e3, 4),
                        "issue_frac": round(issue_s / (adj_avg * 1e-3), 3)}
        else:
            traffic_note = (f"PMC summary is of other kernels or predates the current kernel sources "
                            f"({entry.get('kernels')} against {adj_kernel}): not used")
    kernel_stats = {
        "k_fused_adj": {"launches_per_step": adj_n / args.steps, "avg_ms": round(adj_avg, 4),
                        "ms_per_step": round(adj_ms / args.steps, 2), "bytes_per_launch": int(fused_bytes),
                        "GBps": round(fused_gbs, 1)},
        "k_fused_pass": {"launches_per_step": fwd_n / args.steps, "avg_ms": round(fwd_avg, 4),
                         "ms_per_step": round(fwd_ms / args.steps, 2), "bytes_per_launch": 2 * state_bytes,
                         "GBps": round(2 * state_bytes / (fwd_avg * 1e-3) / 1e9, 1) if fwd_n else None},
    }

    # ---------------------------------------------------------------- batched circuits
    batched = None
    if not args.no_batch:
        batched = bench_batched(N, args, np.complex128)
        batched["complex64"] = bench_batched(N, args, np.complex64)
    serving = None
    if not args.no_serve:
        serving = bench_serving(N, args)
    measurements = None
    if not args.no_measure:
        measurements = bench_measurements(N, args)

    # ---------------------------------------------------------------- 30q single-gate sweep
    sweep = None
    if not args.no_sweep:
        sn = args.sweep_qubits
        sbytes = 16 * (1 << sn)
        log(f"[bench] {sn}q single-gate sweep ({sbytes / 2**30:.0f} GiB state)")
        s = N.State(sn).reset()
        s.apply([("Hadamard", [w], [], False) for w in range(sn)] +
                [("RZ", [w], [0.1 * (w + 1)], False) for w in range(sn)])
        st = StreamTimer(s.stream())
        rows = []
        for t in range(sn):
            op = [("RX", [t], [0.3], False)]
            s.apply(op)  # warm
            ms = st.time(lambda: s.apply(op), reps=args.sweep_reps)
            rows.append({"target": t, "ms": round(ms, 4), "GBps": round(2 * sbytes / (ms * 1e-3) / 1e9, 1)})
        gbs = [r["GBps"] for r in rows]
        # streaming ceiling of the same access (read + write every amplitude in place)
        s.scale(1.0)
        scale_ms = st.time(lambda: s.scale(1.0), reps=args.sweep_reps)
        sweep = {"qubits": sn, "gate": "RX", "bytes_per_pass": 2 * sbytes,
                 "inplace_scale_GBps": round(2 * sbytes / (scale_ms * 1e-3) / 1e9, 1),
                 "min_GBps": min(gbs),
                 "median_GBps": float(np.median(gbs)), "max_GBps": max(gbs),
                 "min_frac_of_peak": round(min(gbs) / HBM_PEAK_GBS, 3), "per_target": rows}
        del s
        N.check(N.lib().lgh_pool_trim(0))

    # ---------------------------------------------------------------- the reference's published workload shape
    published = None
    if not args.no_published:
        published = bench_published(N, ops_mod, args)

    # ---------------------------------------------------------------- CPU oracle baseline (measured components)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, ops, n)
        cpu["config1"]["gpu_s"] = gpu_config1(ops_mod, args)

    rec = {
        "metric": "28q adjoint-Jacobian s",
        "value": round(sec, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (f64)",
        "data": "synthetic (seeded circuit angles)",
        "config": {"workload": f"adjoint Jacobian, {n} qubits c128, {layers} layers RY/RZ + CZ brick, "
                               f"{len(ops)} ops, {n_par} params, <Z0 Z{n - 1}>", "seed": args.seed,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": f"{adj_kernel if adj_kernel.endswith('>') else adj_kernel + ', *>'} "
                                               "(fused adjoint pass: lambda + 1 observable state)",
                     "achieved": round(fused_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fused_gbs / HBM_PEAK_GBS, 3), "traffic": traffic, "traffic_note": traffic_note,
                     "bytes_per_launch": int(fused_bytes), "ms_per_launch": round(adj_avg, 4),
                     "valu_issue": valu},
        "cpu_baseline": cpu,
        "autotune": tune,
        "kernels": kernel_stats,
        "batched": batched,
        "serving": serving,
        "measurements": measurements,
        "gate_sweep": sweep,
        "published_workload": published,
        "jacobian_checksum": float(np.sum(jac)),
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
