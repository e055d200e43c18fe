#!/usr/bin/env python3
"""Batched bordered-banded solve benchmark on one MI355X: mode-solves/s, FP64 complex.

Workload (the microbench configuration of the IMEX hot path): B = 4096 complex128 almost-banded
systems, N = 256, band kl = ku = 4 (N(0,1) + i N(0,1) entries, diagonal + 8), two dense boundary
rows (T_n(1) = 1 and T_n(-1) = (-1)^n), 4 seeded N(0,1) complex right-hand sides per system.
A step is one batched solve of the 4 x 4096 right-hand sides (the solve kernel, factors and data
resident in HBM); the factorization is timed separately.  Accuracy gate before timing: every
system's normwise backward error <= 1e-13.

Prints one JSON line (stdout).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

B, N, KL, KU, R, NRHS = 4096, 256, 4, 4, 2, 4
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def solve_bytes_per_system(n, kl, ku, r, nrhs) -> int:
    """Algorithmic bytes of one system's solve in the stored format: U rows (kl+ku+1) and L rows
    (kl) complex, int32 pivots, V^T and W (r x n), the r x r capacitance LU, and each right-hand
    side read once and written once."""
    return 16 * (2 * kl + ku + 1) * n + 4 * n + 32 * r * n + 16 * r * r + 4 * r + 32 * n * nrhs


def factor_bytes_per_system(n, kl, ku, r) -> int:
    """Algorithmic bytes of one system's factorization: band rows (kl+ku+1) and V^T read; U, L,
    pivots, W (written by the elimination and rewritten by the back substitution, read once) and
    the capacitance LU written."""
    return (16 * (kl + ku + 1) * n + 16 * r * n            # A band + V^T read
            + 16 * (2 * kl + ku + 1) * n + 4 * n           # U, L, pivots written
            + 16 * r * n * 3                               # W: forward write, read + rewrite
            + 16 * r * r + 4 * r + 8)                      # S, pivots, cond


def backward_errors(indptr, indices, vals, x, b):
    rowidx = np.repeat(np.arange(N), np.diff(indptr))
    Ax = np.zeros_like(b)
    np.add.at(Ax.T, rowidx, (vals * x[:, indices]).T)
    nA = np.zeros(b.shape)
    np.add.at(nA.T, rowidx, np.abs(vals).T)
    res = np.abs(Ax - b).max(axis=1)
    return res / (nA.max(axis=1) * np.abs(x).max(axis=1) + np.abs(b).max(axis=1))


def cpu_baseline(indptr, indices, datas, b, seconds: float):
    """The CPU oracle (numpy + LAPACK zgbtrs, the reference's algorithm) on a sample of the same
    systems, one thread; returns mode-solves/s and the sample description."""
    import scipy.sparse as sp
    from oracle.banded import BorderedBanded as OracleBB
    S = 256
    sys_ = [OracleBB(sp.csr_matrix((datas[m], indices, indptr), shape=(N, N)), [0, 1]) for m in range(S)]
    reps, t_total, t0 = 0, 0.0, time.perf_counter()
    while True:
        ts = time.perf_counter()
        for m in range(S):
            sys_[m].solve(b[:, m, :])
        t_total += time.perf_counter() - ts
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    rate = reps * S * NRHS / t_total
    return rate, f"{S} of the {B} systems x {NRHS} rhs, solved {reps} times ({t_total:.1f} s)"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): the "
              "batched solve shards by independent systems with no communication, so the figure at N "
              "GPUs is N times the one-GPU figure by construction and is not measured here.",
              file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    from conatus_amd.runtime.banded import BatchHandle
    from tests.banded_cases import microbench_csr

    rng = np.random.default_rng(0)
    log(f"building {B} systems N={N} kl=ku={KL} r={R}")
    indptr, indices, datas = microbench_csr(rng, B, N, KL, KU, R)
    b = rng.standard_normal((NRHS, B, N)) + 1j * rng.standard_normal((NRHS, B, N))
    log("factoring on the device")
    h = BatchHandle([(N, indptr, indices, d) for d in datas], [0, 1], np.complex128)
    h.reserve(NRHS)
    p = h.props()
    assert (p.kl, p.ku, p.rank) == (KL, KU, R), (p.kl, p.ku, p.rank)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    d = torch.from_numpy(b).cuda()
    h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
    x = d.cpu().numpy()
    vals = np.stack(datas)
    be = max(float(backward_errors(indptr, indices, vals, x[k], b[k]).max()) for k in range(NRHS))
    log(f"accuracy gate: max backward error {be:.2e}")
    if not be <= 1e-13:
        log("accuracy gate FAILED: no timing reported")
        return 1

    # factor timing (k_band_lu + k_border + k_cap_lu from the resident layout)
    for _ in range(args.warmup):
        h.factor(sptr, check=False)
    torch.cuda.synchronize()
    fe = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for e0, e1 in fe:
        e0.record(stream)
        h.factor(sptr, check=False)
        e1.record(stream)
    torch.cuda.synchronize()
    factor_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in fe]))
    log(f"factor: {factor_ms:.4f} ms per batch")

    # solve timing: W warmup + K timed steps, each one launch of the solve kernel on `stream`
    d.copy_(torch.from_numpy(b))
    for _ in range(args.warmup):
        h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record(stream)
        h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
    ms_per_step = wall * 1e3 / args.steps
    mode_solves = B * NRHS / (ms_per_step * 1e-3)
    log(f"solve: {ms_per_step:.4f} ms per step (events {kern_ms:.4f} ms) -> {mode_solves:.4g} mode-solves/s")

    sb = solve_bytes_per_system(N, KL, KU, R, NRHS) * B
    fb = factor_bytes_per_system(N, KL, KU, R) * B
    achieved = sb / (kern_ms * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "k_solve_microbench_traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": "mode-solves/s (batched bordered-banded solve, FP64 complex; IMEX hot path)",
        "value": mode_solves,
        "unit": "mode-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded numpy default_rng(0))",
        "config": {"workload": "batched solve microbench: B=4096 complex128 almost-banded systems, "
                               "N=256, kl=ku=4 (+8 diagonal), 2 dense boundary rows, 4 rhs per system",
                   "batch": B, "n": N, "kl": KL, "ku": KU, "border_rank": R, "nrhs": NRHS},
        "factor_ms": factor_ms,
        "factors_per_s": B / (factor_ms * 1e-3),
        "factor_achieved_GBs": fb / (factor_ms * 1e-3) / 1e9,
        "max_backward_error": be,
        "roofline": {"bound": "hbm", "kernel": "k_solve", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": sb, "launch_ms": kern_ms},
    }
    if not args.no_cpu:
        log(f"cpu baseline (oracle, ~{args.cpu_seconds:.0f} s)")
        rate, sample = cpu_baseline(indptr, indices, datas, b, args.cpu_seconds)
        out["cpu_baseline"] = {"value": rate, "unit": "mode-solves/s", "cores": 1, "kind": "port",
                               "sample": sample}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
