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
    ap.add_argument("--cpu-sample-ops", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1337)
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

    n, layers = args.qubits, args.layers
    ops = ry_rz_cz_brick(n, layers, seed=args.seed)
    n_par = sum(1 for o in ops if o[2])
    state_bytes = 16 * (1 << n)

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

    # ---------------------------------------------------------------- dominant kernel: fused adjoint step
    # Per fused step with one observable state: read+write lambda and the
    # observable state = 4 * state bytes.  Measured as the time difference of
    # two adjoint calls with 2K and K trainable RY ops (fixed costs cancel).
    psi = N.State(n).reset()
    ry_ops = [("RY", [w % n], [0.1 + 0.01 * w], False) for w in range(2 * n)]
    psi.apply(ry_ops)
    timer = StreamTimer(psi.stream())
    bras = [N.State(n)]

    def adj_k(k):
        bras[0].copy_from(psi)
        sub = ry_ops[:k]
        return timer.time(lambda: N.adjoint_jacobian(psi, bras, sub, list(range(k))))

    adj_k(4)
    t1 = min(adj_k(n) for _ in range(2))
    t2 = min(adj_k(2 * n) for _ in range(2))
    fused_ms = (t2 - t1) / n
    fused_bytes = 4 * state_bytes
    fused_gbs = fused_bytes / (fused_ms * 1e-3) / 1e9
    del psi, bras
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile) and n == 28:
        with open(tfile) as f:
            traffic = json.load(f).get("k_adj_dense<double,1>@28q,1obs", {}).get("traffic_bytes")

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

    # ---------------------------------------------------------------- CPU oracle baseline (bounded sample)
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po

        k = args.cpu_sample_ops
        sample_ops = ops[:k]
        log(f"[bench] CPU oracle sample: {n}q, first {k} ops, forward + adjoint")
        t0 = time.perf_counter()
        cpsi = po.apply_ops(n, po.zero_state(n), sample_ops)
        cbra = po.apply_ops(n, cpsi, [("PauliZ", [0], [], False), ("PauliZ", [n - 1], [], False)])
        cjac = po.adjoint_jacobian(n, cpsi, [cbra], sample_ops, list(range(sum(1 for o in sample_ops if o[2]))))
        tcpu = time.perf_counter() - t0
        del cpsi, cbra
        cpu = {"value": round(tcpu * len(ops) / k, 2), "unit": "s", "cores": po.num_threads(), "kind": "port",
               "sample": f"{n}q complex128, first {k} of {len(ops)} ops (forward + adjoint) took {tcpu:.2f} s; "
                         f"scaled by {len(ops)}/{k} ops"}

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
        "roofline": {"bound": "hbm", "kernel": "fused adjoint step (k_adj_dense<double,1>, 1 observable)",
                     "achieved": round(fused_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fused_gbs / HBM_PEAK_GBS, 3), "traffic": traffic,
                     "bytes_per_launch": fused_bytes, "ms_per_launch": round(fused_ms, 4)},
        "cpu_baseline": cpu,
        "gate_sweep": sweep,
        "jacobian_checksum": float(np.sum(jac)),
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
