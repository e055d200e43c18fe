#!/usr/bin/env python3
"""Batched bordered-banded solve benchmark on one MI355X: mode-solves/s, FP64 complex.

Workload (the microbench configuration of the IMEX hot path): B = 4096 complex128 almost-banded
systems, N = 256, band kl = ku = 4 (N(0,1) + i N(0,1) entries, diagonal + 8), two dense boundary
rows (T_n(1) = 1 and T_n(-1) = (-1)^n), 4 seeded N(0,1) complex right-hand sides per system.
A step is one batched solve of the 4 x 4096 right-hand sides (the solve kernel, factors and data
resident in HBM); the factorization is timed separately.  Accuracy gate before timing: every
system's normwise backward error <= 1e-13.  About 0.3 s of untimed solves precede the warm-up
steps so that the timed steps run at the GPU's working clock.  Further legs: IMEX steps/s of 2D Rayleigh-Benard at
128 x 64 (imex_rb2d), an uneven ensemble of 256 RB 2D runs advanced together (ensemble) and a
served stream of RB 2D requests with continuous batching, p50/p99 step latency (served).

Prints one JSON line (stdout).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import gc
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


def cpu_threads() -> int:
    """Host cores the CPU baseline may use: the cores this process may run on, at most 16 (one
    GPU's share of the host on the benchmark box)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _cpu_worker(args):
    """One CPU-baseline worker (a fresh interpreter, no GPU): factor its systems with the oracle,
    wait for the others, then solve them repeatedly for `seconds`; returns (right-hand sides
    solved, seconds spent solving)."""
    indptr, indices, datas, b, seconds, barrier = args
    import scipy.sparse as sp
    from threadpoolctl import threadpool_limits
    from oracle.banded import BorderedBanded as OracleBB
    n = len(indptr) - 1
    with threadpool_limits(limits=1):
        sys_ = [OracleBB(sp.csr_matrix((d, indices, indptr), shape=(n, n)), [0, 1]) for d in datas]
        barrier.wait()
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            for m, o in enumerate(sys_):
                o.solve(b[:, m, :])
            done += len(sys_) * b.shape[0]
        return done, time.perf_counter() - t0


def cpu_baseline(indptr, indices, datas, b, seconds: float):
    """The CPU oracle (numpy + LAPACK zgbtrs, the reference's algorithm) on a sample of the same
    systems, split over workers by the reference's batch rule t = min(allowed, B, isqrt(B n / 900))
    (python/conatus/runtime/banded.py:28-33, B = the whole batch).  The reference's workers are C++
    threads; the oracle is Python, whose threads would serialise on the interpreter lock, so each
    worker is a process of its own (spawned: no GPU state is inherited), single-threaded BLAS, all
    solving at once after a common barrier.  Returns mode-solves/s, the sample description and the
    worker count."""
    import multiprocessing as mp
    from oracle.banded import threads_for
    S = 1024
    t = threads_for(B, N, cpu_threads())
    chunks = [range(q * S // t, (q + 1) * S // t) for q in range(t)]
    ctx = mp.get_context("spawn")
    with ctx.Manager() as man:
        barrier = man.Barrier(t)
        jobs = [(indptr, indices, [datas[m] for m in c], np.ascontiguousarray(b[:, c.start:c.stop, :]), seconds, barrier)
                for c in chunks]
        with ctx.Pool(t) as pool:
            res = pool.map(_cpu_worker, jobs)
    done = sum(r[0] for r in res)
    span = max(r[1] for r in res)
    rate = done / span
    return rate, (f"{S} of the {B} systems x {NRHS} rhs over {t} worker processes (min(allowed {cpu_threads()}, "
                  f"B, isqrt(B n / 900))), {done} right-hand sides in {span:.1f} s"), t


def quiet_heap() -> None:
    """Collect the garbage of the set-up and move the survivors out of the collector's view, so
    that no full collection pauses the host inside a timed region (the timed regions submit device
    work from a Python loop).  Called before a leg's warm-up, not between it and the timed region:
    the collection idles the GPU long enough for its clock to drop (the first ~100 microbench
    solves after a 0.5 s pause ran up to 10 % slower)."""
    gc.collect()
    gc.freeze()


def rb2d_leg(steps: int, warmup: int, cpu_seconds: float, with_cpu: bool) -> dict:
    """IMEX steps/s of 2D Rayleigh-Benard at 128 x 64 (Ra = 1e5, Pr = 1, walls at z = -1, 1,
    SBDF2): 64 complex mode systems (n = 263, kl = 12, ku = 18, r = 7) per step.  Accuracy gate
    first (20 steps at dt = 1e-3 against the CPU oracle, <= 1e-11 relative); the timed steps are
    graph replays of full SBDF2 steps at dt = 1e-4, where the seeded flow stays resolved for the
    200 steps (dt = 1e-3 exceeds the explicit-advection limit after about 60 steps, in the oracle
    too; the cost of a step does not depend on dt)."""
    import torch
    from conatus_amd.problems import RayleighBenard2D
    from conatus_amd.timestep import IVPSolver
    from oracle.rb2d import RB2DOracle
    pb = RayleighBenard2D(Nx=128, Nz=64, Lx=4.0, Ra=1e5, Pr=1.0)
    X0 = pb.initial_state(seed=0)
    g = IVPSolver(pb, "SBDF2", 1e-3)
    g.set_state(X0)
    g.run(20)
    ref = RB2DOracle(pb, "SBDF2", 1e-3).run(X0, 20)
    gate = float(np.max(abs(g.state() - ref)) / np.max(abs(ref)))
    log(f"rb2d accuracy gate: relative difference after 20 steps {gate:.2e}")
    if not gate <= 1e-11:
        raise RuntimeError(f"rb2d accuracy gate failed ({gate:.3e})")
    s = IVPSolver(pb, "SBDF2", 1e-4)
    quiet_heap()
    # warm-up: replays from the initial state for about half a second (the GPU clock ramps up
    # over the first ~0.1 s of work), both ring phases captured
    tw = time.perf_counter()
    while True:
        s.set_state(X0)
        s.run_graph(max(warmup, 100))
        s.run_graph(16)                  # a continuation call, as the timed one is
        torch.cuda.synchronize()
        if time.perf_counter() - tw > 0.5:
            break
    s.set_state(X0)
    s.run_graph(4)                       # startup order 1, then full-order graph replays
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    s.run_graph(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    X = s.state()
    if not np.all(np.isfinite(X)):
        raise RuntimeError("rb2d state not finite after the timed steps")
    out = {"workload": "RB 2D 128x64 SBDF2 (64 complex mode systems n=263, kl=12, ku=18, r=7)",
           "steps": steps, "ms_per_step": ms, "steps_per_s": 1e3 / ms,
           "wall_ms_per_step": wall * 1e3 / steps, "mode_solves_per_s": pb.K * 1e3 / ms,
           "gate_relative_difference_20_steps": gate, "graph": True}
    # the same problem stepped with RK443 (four stages, four solves and three explicit terms per
    # step: the scheme of the reference's published GPU timings at this size)
    rk = IVPSolver(pb, "RK443", 1e-4)
    quiet_heap()
    tw = time.perf_counter()
    while time.perf_counter() - tw < 0.3:
        rk.set_state(X0)
        rk.run_graph(32)
        rk.run_graph(8)
        torch.cuda.synchronize()
    rk.set_state(X0)
    rk.run_graph(8)
    torch.cuda.synchronize()
    e0.record(stream)
    rk.run_graph(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    if not np.all(np.isfinite(rk.state())):
        raise RuntimeError("rb2d RK443 state not finite after the timed steps")
    out["rk443_ms_per_step"] = e0.elapsed_time(e1) / steps
    if with_cpu:
        # blocks of 100 steps from the initial state (the flow stays resolved), factorizations of
        # both startup orders built before the timing
        o = RB2DOracle(pb, "SBDF2", 1e-4)
        o.run(X0, 3)
        n_done, t_cpu, tc0 = 0, 0.0, time.perf_counter()
        while time.perf_counter() - tc0 < cpu_seconds:
            o.MXh, o.Fh = [], []
            ts = time.perf_counter()
            Xc = o.run(X0, 100)
            t_cpu += time.perf_counter() - ts
            n_done += 100
        if not np.all(np.isfinite(Xc)):
            raise RuntimeError("rb2d CPU baseline state not finite")
        out["cpu_baseline"] = {"value": n_done / t_cpu, "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": f"{n_done} SBDF2 steps of the same problem in blocks of 100 from the "
                                         f"initial state ({t_cpu:.1f} s, numpy/pocketfft + LAPACK zgbtrs)"}
    return out


def channel_leg(steps: int, gate_steps: int = 10) -> dict:
    """IMEX steps/s of the 3D plane channel at 128 x 97 x 128 (Lx = 4 pi, Lz = 2 pi, Re = 2000,
    laminar Poiseuille flow + a seeded divergence-free perturbation of kinetic energy 1e-4,
    SBDF3, dt = 5e-4): 8128 complex mode systems (n = 395, r = 7) per step and the explicit term
    on the 192 x 145 x 192 grid.  Accuracy gate first, at 16 x 33 x 16 against the CPU oracle
    (velocity <= 2e-10 relative after `gate_steps` steps, the whole state reported: two
    backward-stable CPU solves of this case already differ by 3e-11 in the velocity and 3e-8 in
    the tau multipliers, the mode systems' condition numbers being near 1e6; the oracle needs
    minutes per step at full size);
    at full size the state must stay finite and the perturbation energy bounded.  The timed steps
    are graph replays of full-order steps."""
    import torch
    from conatus_amd.problems import Channel3D
    from conatus_amd.timestep import IVPSolver
    from oracle.channel3d import Channel3DOracle
    small = Channel3D(Nx=16, Ny=33, Nz=16, Re=2000.0)
    Xs = small.initial_state(seed=1, energy=1e-3)
    g = IVPSolver(small, "SBDF3", 5e-4)
    g.set_state(Xs)
    g.run(gate_steps)
    ref = Channel3DOracle(small, "SBDF3", 5e-4).run(Xs, gate_steps)
    vel = np.concatenate([small.field_index(f) for f in (0, 1, 2)])
    got = g.state()
    gate = float(np.max(abs(got[:, vel] - ref[:, vel])) / np.max(abs(ref[:, vel])))
    full = float(np.max(abs(got - ref)) / np.max(abs(ref)))
    log(f"channel accuracy gate (16x33x16): velocity relative difference after {gate_steps} steps {gate:.2e} "
        f"(whole state incl. pressure and taus {full:.2e})")
    if not gate <= 2e-10:                # the rounding floor of this case is 3e-11 (two backward-stable solves)
        raise RuntimeError(f"channel accuracy gate failed ({gate:.3e})")
    del g
    pb = Channel3D(Nx=128, Ny=97, Nz=128, Re=2000.0)
    X0 = pb.initial_state(seed=0, energy=1e-4)
    s = IVPSolver(pb, "SBDF3", 5e-4)
    log("channel: " + s.explain())
    quiet_heap()
    s.set_state(X0)
    s.run_graph(48)                      # warm-up, about 0.1 s
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    s.run_graph(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    X = s.state()
    if not np.all(np.isfinite(X)):
        raise RuntimeError("channel state not finite after the timed steps")
    energy = pb.kinetic_energy(pb.velocity(X - pb.poiseuille()))
    if not 0 < energy < 1e-3:
        raise RuntimeError(f"channel perturbation energy {energy:.3e} out of range")
    return {"workload": "channel 3D 128x97x128 SBDF3 (8128 complex mode systems n=395, r=7; grid 192x145x192)",
            "steps": steps, "ms_per_step": ms, "steps_per_s": 1e3 / ms, "wall_ms_per_step": wall * 1e3 / steps,
            "mode_solves_per_s": pb.K * 1e3 / ms, "gate_velocity_relative_difference": gate,
            "gate_state_relative_difference": full,
            "gate_workload": f"16x33x16, {gate_steps} SBDF3 steps at dt = 5e-4",
            "perturbation_energy_after": energy, "graph": True}


def channel_pt_step_bytes(pb) -> dict:
    """Algorithmic HBM bytes of one SBDF3 step of the poloidal-toroidal channel, by kernel: every
    launch reads its inputs and writes its outputs once (state S = 4 P (nP + nS) doubles).
    Explicit term: A reads the state's coefficients and writes u, v, w at the nodes (Kx My 3 Kz
    complex); B reads them and writes 3 z-lines of Mz per (kx, y_j); C reads those and writes 6
    products per (kx, y_j, z_m); D reads them and writes 6 x Kz per (kx, y_j); E reads those and
    writes F.  Stage: X, three F levels and two M X levels read, M X and the right-hand side (over
    X) written.  Solves (real factors, four right-hand sides each): per system
    8 (2 kl + ku + 1) n + 4 n + 16 r n + 8 r^2 + 16 n nrhs."""
    S = 8 * 4 * pb.P * (pb.nP + pb.nS)
    z = 16
    h1 = z * pb.Kx * pb.My * 3 * pb.Kz
    h2 = z * pb.My * 3 * pb.Kx * pb.Mz
    h3 = z * pb.My * 6 * pb.Kx * pb.Mz
    h4 = z * pb.Kx * pb.My * 6 * pb.Kz
    solve = lambda n, kl, ku, r: pb.P * (8 * (2 * kl + ku + 1) * n + 4 * n + 16 * r * n + 8 * r * r + 16 * n * 4)  # noqa: E731
    out = {"yback": S + h1, "zback": h1 + h2, "xline": h2 + h3, "zfwd": h3 + h4, "yfwd": h4 + S,
           "stage": 6 * S + 2 * S, "solve_phi": solve(pb.nP, 4, 4, 4), "solve_psi": solve(pb.nS, 2, 2, 2)}
    out["total"] = float(sum(out.values()))
    return out


def channel_pt_leg(steps: int) -> dict:
    """IMEX steps/s of the 3D plane channel in the configuration's velocity-vorticity
    (poloidal-toroidal) form at 128 x 97 x 128 (Lx = 4 pi, Lz = 2 pi, Re = 2000, Poiseuille + a
    seeded divergence-free perturbation of energy 1e-4, SBDF3, dt = 5e-4): 8128 complex modes as
    4096 real phi systems (n = 101, kl = ku = 4, r = 4) and 4096 real psi systems (n = 99,
    kl = ku = 2, r = 2), four right-hand sides each, and the explicit term from six products on
    the 192 x 145 x 192 grid.  Accuracy gate first: 20 SBDF3 steps at 16 x 17 x 12 against the CPU
    oracle (<= 1e-11 relative; the 128 x 97 x 128 parity is a GPU test).  The timed steps are graph
    replays of full-order steps."""
    import torch
    from conatus_amd.channel_pt import ChannelPT3D, ChannelPTSolver
    from oracle.channel_pt import ChannelPTOracle
    small = ChannelPT3D(Nx=16, Ny=17, Nz=12, Re=800.0)
    Xs = small.initial_state(seed=3, energy=1e-3, kmax=3)
    g = ChannelPTSolver(small, "SBDF3", 2e-3)
    g.set_state(Xs)
    g.run(20)
    ref = ChannelPTOracle(16, 17, 12, 800.0, small.Lx, small.Lz, small.G, scheme="SBDF3", dt=2e-3).run(Xs, 20)
    gate = float(np.max(abs(g.state() - ref)) / np.max(abs(ref)))
    log(f"channel (v-omega) accuracy gate (16x17x12, 20 SBDF3 steps): relative difference {gate:.2e}")
    if not gate <= 1e-11:
        raise RuntimeError(f"channel (v-omega) accuracy gate failed ({gate:.3e})")
    pb = ChannelPT3D(Nx=128, Ny=97, Nz=128, Re=2000.0)
    X0 = pb.initial_state(seed=0, energy=1e-4)
    s = ChannelPTSolver(pb, "SBDF3", 5e-4)
    log("channel (v-omega): " + s.explain())
    quiet_heap()
    s.set_state(X0)
    s.run_graph(192)                     # warm-up, about 0.1 s
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    s.run_graph(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    X = s.state()
    if not np.all(np.isfinite(X)):
        raise RuntimeError("channel (v-omega) state not finite after the timed steps")
    energy = pb.kinetic_energy(pb.velocity(X - pb.poiseuille()))
    if not 0 < energy < 1e-3:
        raise RuntimeError(f"channel (v-omega) perturbation energy {energy:.3e} out of range")
    nb = channel_pt_step_bytes(pb)
    ach = nb["total"] / (ms * 1e-3) / 1e9
    survey = 1.70e9            # the configuration's v-omega step model (solve, history, M X, F traffic)
    return {"workload": "channel 3D 128x97x128 velocity-vorticity (poloidal-toroidal) SBDF3, dt = 5e-4: 4096 real "
                        "phi systems (n=101, kl=ku=4, r=4) + 4096 real psi systems (n=99, kl=ku=2, r=2), 4 rhs "
                        "each (8128 complex modes); grid 192x145x192",
            "steps": steps, "ms_per_step": ms, "steps_per_s": 1e3 / ms, "wall_ms_per_step": wall * 1e3 / steps,
            "mode_solves_per_s": pb.K * 1e3 / ms, "gate_relative_difference": gate,
            "gate_workload": "16x17x12, 20 SBDF3 steps at dt = 2e-3 against the CPU oracle",
            "perturbation_energy_after": energy, "graph": True,
            "algorithmic_bytes_per_step": nb, "achieved_GBs": ach, "roofline_frac": ach / HBM_PEAK_GBS,
            "survey_model_bytes_per_step": survey,
            "survey_model_roofline_frac": survey / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def rb2d_step_bytes(pb) -> float:
    """Algorithmic HBM bytes of one SBDF2 step of a RB 2D run: the batched solve of its K mode
    systems (band rows of kl + ku + 1 = 31 entries for U and 12 for L, pivots, V^T and W of r = 7
    rows, S, the right-hand side read and x written: 16 (43 n + 2 r n + 2 n) + 4 n + 16 r^2 per
    system), the explicit term (X read, the 3 K Mz line products written and read, F rows
    written: 16 K (2 n + 6 Mz)) and the stage (history M X and F of two levels read, M X and the
    right-hand side written, CSR of M with ~3 entries per row: 16 K n (4 + 2) + 12 * 3 K n)."""
    K, n, Mz = pb.K, pb.n, pb.Mz
    kl, ku, r = 12, 18, 7
    solve = K * (16 * ((kl + ku + 1 + kl) * n + 2 * r * n + 2 * n) + 4 * n + 16 * r * r)
    explicit = 16 * K * (2 * n + 6 * Mz)
    stage = 16 * K * n * 6 + 12 * 3 * K * n
    return float(solve + explicit + stage)


def single_run_ms(pb, dt: float, steps: int = 40) -> float:
    """GPU time per step of one run stepped alone (IVPSolver, graph replays, after warm-up)."""
    import torch
    from conatus_amd.timestep import IVPSolver
    s = IVPSolver(pb, "SBDF2", dt)
    X0 = pb.initial_state(seed=0)
    quiet_heap()
    s.set_state(X0)
    s.run_graph(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run_graph(steps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def ensemble_leg(runs: int, steps: int, dt: float) -> dict:
    """Uneven ensemble: `runs` independent RB 2D runs, Ny drawn from {32, 48, 64, 96, 128} (seeded),
    Nx = 2 Ny, Ra log-uniform in [1e4, 1e6], Pr = 1, seeded noise initial states as the 128 x 64
    leg, SBDF2; all runs advanced together by one EnsembleSet (one batched step per size class,
    all classes in one captured graph, kernel mode per class by the autotuner).  Reported beside
    per-run sequential GPU stepping: each size class's single-run step time (measured) times its
    runs, summed."""
    import torch
    from conatus_amd.ensembles import EnsembleSet
    from conatus_amd.problems import RayleighBenard2D
    rng = np.random.default_rng(3)
    nys = rng.choice([32, 48, 64, 96, 128], size=runs)
    ras = 10 ** rng.uniform(4, 6, size=runs)
    t0 = time.perf_counter()
    pbs = [RayleighBenard2D(Nx=2 * int(ny), Nz=int(ny), Lx=4.0, Ra=float(ra), Pr=1.0) for ny, ra in zip(nys, ras)]
    X0 = [pb.initial_state(seed=1000 + i) for i, pb in enumerate(pbs)]
    es = EnsembleSet(pbs, "SBDF2", dt)
    es.set_states(X0)
    log(f"ensemble set up in {time.perf_counter() - t0:.1f} s: {es.explain()}")
    quiet_heap()
    es.run_graph(3)                        # startup, capture
    es.run_graph(20)                       # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    es.run_graph(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    got = es.states()
    if not all(np.all(np.isfinite(x)) for x in got):
        raise RuntimeError("ensemble state not finite after the timed steps")
    # parity gate at this workload: the first run of every size class against the same run
    # stepped alone (bit for bit) and against the CPU oracle (relative difference reported)
    from conatus_amd.ensembles import discretization_key
    from conatus_amd.timestep import IVPSolver
    from oracle.rb2d import RB2DOracle
    total = 3 + 20 + steps
    first: dict = {}
    for i, pb in enumerate(pbs):
        first.setdefault(discretization_key(pb), i)
    gate = {}
    for key, i in sorted(first.items(), key=lambda kv: pbs[kv[1]].Nz):
        pb = pbs[i]
        gi, _ = es.where[i]
        one = IVPSolver(pb, "SBDF2", dt, mode=es.groups[gi].mode)
        one.set_state(X0[i])
        one.run_graph(3)
        one.run_graph(20)
        one.run_graph(steps)
        if not np.array_equal(got[i], one.state()):
            raise RuntimeError(f"ensemble run {i} ({pb.Nx}x{pb.Nz}) differs from the same run stepped alone")
        ref = RB2DOracle(pb, "SBDF2", dt).run(X0[i], total)
        r = float(np.max(abs(got[i] - ref)) / np.max(abs(ref)))
        if not r <= 1e-10:
            raise RuntimeError(f"ensemble run {i} ({pb.Nx}x{pb.Nz}): oracle difference {r:.2e}")
        gate[f"{pb.Nx}x{pb.Nz}"] = r
    log("ensemble gate: one run per class bit for bit as stepped alone; oracle differences after "
        f"{total} steps " + ", ".join(f"{k} {v:.1e}" for k, v in gate.items()))
    per_class = {}
    for ny in sorted(set(int(v) for v in nys)):
        pb = RayleighBenard2D(Nx=2 * ny, Nz=ny, Lx=4.0, Ra=1e5, Pr=1.0)
        per_class[ny] = single_run_ms(pb, dt)
    seq_ms = sum(per_class[int(ny)] for ny in nys)          # one step of every run, one run after another
    ms = wall * 1e3 / steps
    modes = sum(pb.K for pb in pbs)
    nbytes = sum(rb2d_step_bytes(pb) for pb in pbs)
    return {"workload": f"{runs} RB 2D runs, Ny in {{32,48,64,96,128}} (seeded), Nx = 2 Ny, Ra log-uniform "
                        f"[1e4, 1e6], Pr = 1, SBDF2 dt = {dt:g}, {steps} timed steps after 3",
            "runs": runs, "size_classes": {str(k): int((nys == k).sum()) for k in sorted(set(int(v) for v in nys))},
            "ms_per_ensemble_step": ms, "run_steps_per_s": runs * 1e3 / ms,
            "mode_solves_per_s": modes * 1e3 / ms,
            "achieved_GBs": nbytes / (ms * 1e-3) / 1e9, "roofline_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes_per_step": nbytes,
            "sequential_single_run_ms_per_ensemble_step": seq_ms,
            "speedup_vs_sequential_single_runs": seq_ms / ms,
            "single_run_ms_per_step": {str(k): v for k, v in per_class.items()},
            "gate": {"bitwise_vs_single_runs": True, "steps": total, "oracle_relative_difference": gate},
            "tuning": es.tuning}


def served_leg(nreq: int, dt: float, channel_share: float = 0.1) -> dict:
    """Served stream: Poisson arrivals (seeded, mean 50 requests/s) mixing RB 2D runs of the
    128 x 64 leg's physics (Ra = 1e5, Pr = 1) with Ny drawn from {32, 48, 64, 96, 128}, Nx = 2 Ny,
    seeded noise initial states, and (a seeded tenth of the requests) 3D channel runs at
    64 x 65 x 64 (Re = 2000, Lx = 4 pi, Lz = 2 pi, Poiseuille + perturbation of energy 1e-4);
    50-500 SBDF2 steps each, admitted by HBM footprint up to 200 GB and continuously batched
    (conatus_amd.serving).  Reports p50/p99 of the per-step latency every run saw and the
    aggregate run-steps/s, beside per-run sequential GPU stepping of the same requests
    (single-run step times per size class times steps)."""
    from conatus_amd.problems import Channel3D, RayleighBenard2D
    from conatus_amd.serving import Request, Server
    rng = np.random.default_rng(4)
    gaps = rng.exponential(1 / 50.0, size=nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    nys = rng.choice([32, 48, 64, 96, 128], size=nreq)
    steps = rng.integers(50, 501, size=nreq)
    is_channel = rng.random(nreq) < channel_share

    def problem(cls):
        if cls == "channel":
            return Channel3D(Nx=64, Ny=65, Nz=64, Re=2000.0, Lx=4 * np.pi, Lz=2 * np.pi)
        return RayleighBenard2D(Nx=2 * cls, Nz=cls, Lx=4.0, Ra=1e5, Pr=1.0)

    classes = ["channel" if is_channel[i] else int(nys[i]) for i in range(nreq)]
    templates = {c: problem(c) for c in sorted(set(classes), key=str)}
    reqs = []
    for i in range(nreq):
        pb = templates[classes[i]]
        reqs.append(Request(pb, int(steps[i]), pb.initial_state(seed=2000 + i), arrival=float(arrivals[i])))
    srv = Server("SBDF2", dt, hbm_budget=200e9, page_size=8, channel_page_size=2)
    # every size class tuned, its pages and the single-run factors built before the clock starts
    rb = [pb for c, pb in templates.items() if c != "channel"]
    srv.prepare(rb, pages=3)
    if "channel" in templates:
        srv.prepare([templates["channel"]], pages=8, startup_pages=3)
    built = srv.pages_built
    quiet_heap()
    st = srv.serve(reqs, clock="wall")
    bad = sum(not np.all(np.isfinite(q.result)) for q in reqs)
    if bad:
        raise RuntimeError(f"served leg: {bad} results not finite")
    per_class = {c: single_run_ms(pb, dt) for c, pb in templates.items()}
    seq_s = sum(per_class[classes[i]] * int(steps[i]) for i in range(nreq)) * 1e-3
    st.update({"workload": f"{nreq} requests, Poisson arrivals mean 50/s (seeded): RB 2D Ra = 1e5, Pr = 1, "
                           f"Ny in {{32..128}}, Nx = 2 Ny, and {int(is_channel.sum())} 3D channel runs 64x65x64 "
                           f"(Re = 2000); 50-500 SBDF2 steps each, dt = {dt:g}, HBM budget 200 GB",
               "arrival_span_s": float(arrivals[-1]),
               "channel_requests": int(is_channel.sum()),
               "pages_built_while_serving": srv.pages_built - built,
               "sequential_single_run_steps_per_s": st["steps"] / seq_s,
               "speedup_vs_sequential_single_runs": (st["steps"] / st["wall_s"]) / (st["steps"] / seq_s),
               "single_run_ms_per_step": {str(k): v for k, v in per_class.items()}})
    return st


def profile_summary(path: str) -> str:
    """Markdown table of a rocprofv3 kernel trace (``--kernel-trace`` database, ``*_results.db``
    under ``path``): per kernel name the dispatch count, the average and total duration and the
    share of the summed kernel time, longest total first."""
    import glob
    import sqlite3
    dbs = sorted(glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True))
    if not dbs:
        raise SystemExit(f"no rocprofv3 results database under {path}")
    rows = []
    for db in dbs:
        con = sqlite3.connect(db)
        rows += con.execute("select name, count(*), sum(end - start) from kernels group by name").fetchall()
        con.close()
    agg: dict = {}
    for name, n, tot in rows:
        a = agg.setdefault(name, [0, 0])
        a[0] += n
        a[1] += tot
    total = sum(v[1] for v in agg.values()) or 1
    out = ["| kernel | calls | avg µs | total µs | share |", "|---|---|---|---|---|"]
    for name, (n, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        short = name if len(name) <= 90 else name[:87] + "..."
        out.append(f"| `{short}` | {n} | {tot / n / 1e3:.1f} | {tot / 1e3:.1f} | {100 * tot / total:.1f} % |")
    return "\n".join(out)


def traffic_summary(path: str) -> str:
    """HBM bytes per launch of the microbench's solve kernel (k_solve_s) from four rocprofv3
    ``--pmc ... -f csv`` runs under ``path``: ``prof_fetch`` and ``prof_write`` (FETCH_SIZE and
    WRITE_SIZE of ``bench.py --no-rb2d ...``), ``cal_fetch`` and ``cal_write`` (the same counters of
    ``bench.py --calibrate``, whose multiply kernel reads and writes 512 MiB per launch).  Each
    counter is scaled by expected / measured KiB of the calibration kernel.  Prints the JSON record
    that the bench line's ``roofline.traffic`` reads."""
    import csv
    import glob

    def per_dispatch(sub: str, counter: str, want) -> list:
        vals: dict = {}
        for f in sorted(glob.glob(os.path.join(path, sub, "**", "*counter_collection.csv"), recursive=True)):
            with open(f, newline="") as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter or not want(row.get("Kernel_Name", "")):
                        continue
                    key = (f, row.get("Dispatch_Id"))
                    vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
        if not vals:
            raise SystemExit(f"no {counter} rows for the kernel under {os.path.join(path, sub)}")
        return list(vals.values())

    expected_kib = (1 << 26) * 8 / 1024
    is_cal = lambda name: "Mul" in name          # the calibration's y = 2 x kernel (not torch's fill of x)
    is_solve = lambda name: "k_solve_s" in name
    fs = expected_kib / float(np.mean(per_dispatch("cal_fetch", "FETCH_SIZE", is_cal)))
    ws = expected_kib / float(np.mean(per_dispatch("cal_write", "WRITE_SIZE", is_cal)))
    names = set()
    for f in glob.glob(os.path.join(path, "prof_fetch", "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            names |= {r.get("Kernel_Name", "") for r in csv.DictReader(fh) if is_solve(r.get("Kernel_Name", ""))}
    rd = float(np.mean(per_dispatch("prof_fetch", "FETCH_SIZE", is_solve))) * 1024 * fs
    wr = float(np.mean(per_dispatch("prof_write", "WRITE_SIZE", is_solve))) * 1024 * ws
    return json.dumps({"kernel": sorted(names)[0] if names else "k_solve_s", "hbm_bytes_per_launch": rd + wr,
                       "read_bytes": rd, "write_bytes": wr, "fetch_scale": fs, "write_scale": ws,
                       "source": "profiles/k_solve_microbench.md"}, indent=1)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-rb2d", action="store_true", help="skip the RB 2D IMEX steps/s leg")
    ap.add_argument("--rb2d-steps", type=int, default=200)
    ap.add_argument("--no-channel", action="store_true", help="skip the 3D channel IMEX leg (primitive variables)")
    ap.add_argument("--no-channel-pt", action="store_true",
                    help="skip the 3D channel IMEX leg in the velocity-vorticity form")
    ap.add_argument("--channel-steps", type=int, default=50)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the uneven-ensemble leg")
    ap.add_argument("--ensemble-runs", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--served-requests", type=int, default=120)
    ap.add_argument("--only-channel-pt", action="store_true",
                    help="run only the velocity-vorticity channel leg (--channel-steps timed steps) and print its "
                         "JSON object: the command the channel's rocprofv3 summaries are taken from")
    ap.add_argument("--profile-summary", metavar="DIR",
                    help="print the per-kernel table (calls, average and total duration, share) of the rocprofv3 "
                         "kernel-trace database under DIR and exit (no GPU needed)")
    ap.add_argument("--traffic-summary", metavar="DIR",
                    help="print the HBM traffic record of the solve kernel from the rocprofv3 --pmc runs under DIR "
                         "(prof_fetch, prof_write, cal_fetch, cal_write) and exit (no GPU needed)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a plain streaming kernel y = 2 x over 2^26 doubles (512 MiB read and 512 MiB "
                         "written per launch, 5 launches): the reference reading for the HBM byte counters")
    args = ap.parse_args()
    if args.profile_summary:
        print(profile_summary(args.profile_summary))
        return 0
    if args.traffic_summary:
        print(traffic_summary(args.traffic_summary))
        return 0
    # The served leg steps up to ~30 pages on their own streams; HIP maps streams onto
    # GPU_MAX_HW_QUEUES hardware queues (4 by default), so that a small page's step can wait behind
    # a large one in a shared queue (per-class p50 2-3 ms at 4 queues, 0.3-1.4 ms at 16).  Set
    # before the first HIP call (raising a smaller value from the environment, 4 on the GPU box);
    # the single-stream legs are unaffected.
    try:
        hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        hwq = 4
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(hwq, 16), 32))
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
    if args.calibrate:
        xs = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
        ys = torch.empty_like(xs)
        for _ in range(5):
            torch.mul(xs, 2.0, out=ys)
        torch.cuda.synchronize()
        log("calibration kernel done")
        return 0
    if args.only_channel_pt:
        log(f"3D channel (velocity-vorticity) IMEX leg ({args.channel_steps} graph-replayed SBDF3 steps)")
        c = channel_pt_leg(args.channel_steps)
        log(f"channel (v-omega): {c['ms_per_step']:.4f} ms per SBDF3 step, {c['achieved_GBs']:.0f} GB/s "
            f"({100 * c['roofline_frac']:.0f} % of HBM)")
        print(json.dumps(c), flush=True)
        return 0
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

    # clock warm-up (about 0.3 s of solves: the GPU clock ramps up over the first ~0.1 s of work)
    quiet_heap()
    tw = time.perf_counter()
    while time.perf_counter() - tw < 0.3:
        for _ in range(20):
            h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
        torch.cuda.synchronize()

    # factor timing (k_band_lu, border completion k_border_s, k_cap_lu from the resident layout)
    for _ in range(args.warmup):
        h.factor(sptr, check=False)
    torch.cuda.synchronize()
    f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    f0.record(stream)
    for _ in range(args.steps):
        h.factor(sptr, check=False)
    f1.record(stream)
    torch.cuda.synchronize()
    factor_ms = f0.elapsed_time(f1) / args.steps
    log(f"factor: {factor_ms:.4f} ms per batch")

    # solve timing: W warmup + K timed steps, each one launch of the solve kernel on `stream`,
    # back to back; HIP events on that stream at both ends of the timed region give the average
    # launch duration (no commands between the launches)
    tw = time.perf_counter()               # the clock back up after the factorizations
    while time.perf_counter() - tw < 0.1:
        for _ in range(20):
            h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
        torch.cuda.synchronize()
    d.copy_(torch.from_numpy(b))
    for _ in range(args.warmup):
        h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        h.solve_ptr(d.data_ptr(), NRHS, B * N, False, sptr)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = e0.elapsed_time(e1) / args.steps
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
        rate, sample, threads = cpu_baseline(indptr, indices, datas, b, args.cpu_seconds)
        out["cpu_baseline"] = {"value": rate, "unit": "mode-solves/s", "cores": threads, "kind": "port",
                               "sample": sample}
    else:
        out["cpu_baseline"] = None
    if not args.no_rb2d:
        log(f"RB 2D IMEX leg ({args.rb2d_steps} graph-replayed SBDF2 steps)")
        out["imex_rb2d"] = rb2d_leg(args.rb2d_steps, args.warmup, min(args.cpu_seconds, 5.0), not args.no_cpu)
        log(f"rb2d: {out['imex_rb2d']['ms_per_step']:.4f} ms per SBDF2 step, "
            f"{out['imex_rb2d']['rk443_ms_per_step']:.4f} ms per RK443 step")
    if not args.no_channel:
        log(f"3D channel IMEX leg ({args.channel_steps} graph-replayed SBDF3 steps)")
        out["imex_channel"] = channel_leg(args.channel_steps)
        log(f"channel: {out['imex_channel']['ms_per_step']:.3f} ms per SBDF3 step")
    if not args.no_channel_pt:
        log(f"3D channel (velocity-vorticity) IMEX leg ({args.channel_steps} graph-replayed SBDF3 steps)")
        out["imex_channel_pt"] = channel_pt_leg(max(args.channel_steps, 100))
        c = out["imex_channel_pt"]
        log(f"channel (v-omega): {c['ms_per_step']:.3f} ms per SBDF3 step, {c['achieved_GBs']:.0f} GB/s "
            f"({100 * c['roofline_frac']:.0f} % of HBM)")
    if not args.no_ensemble:
        log(f"uneven ensemble leg ({args.ensemble_runs} runs)")
        out["ensemble"] = ensemble_leg(args.ensemble_runs, args.ensemble_steps, 1e-4)
        e = out["ensemble"]
        log(f"ensemble: {e['ms_per_ensemble_step']:.3f} ms per step of all runs, {e['run_steps_per_s']:.4g} "
            f"run-steps/s, {e['speedup_vs_sequential_single_runs']:.1f}x per-run sequential")
    if not args.no_served:
        log(f"served leg ({args.served_requests} requests)")
        out["served"] = served_leg(args.served_requests, 1e-4)
        v = out["served"]
        log(f"served: p50 {v['p50_step_ms']:.3f} ms, p99 {v['p99_step_ms']:.3f} ms per step, "
            f"{v['steps_per_s']:.4g} run-steps/s, {v['speedup_vs_sequential_single_runs']:.1f}x per-run sequential")
        log("served per class (p50 / p99 ms per step): " + ", ".join(
            f"{g} {c['p50_step_ms']:.2f} / {c['p99_step_ms']:.2f}" for g, c in v["per_class"].items()))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
