"""Device-memory plan of an IVP run (the reference's footprint planner, runtime/memory.py:1-62).

The reference replays the allocation program of its CUDA backend through a model of CuPy's pool,
because that backend allocates and frees temporaries every step.  Here every device buffer of a
run is allocated once, before its first step, and a step allocates nothing (HIP graphs replay into
the same buffers), so the plan is exact bookkeeping: one line per buffer, by category, with the
sizes the library and the Python layers allocate:

* state and the continuation copy of the returned state;
* integrator buffers: SBDF history levels (M X, F) and the new F, RK stage states / F / L X, CN
  levels (M X, F, L X), the right-hand side;
* CSR operators M and L (values, column indices, row pointers);
* factor handles, one per implicit operator alpha M + beta L the scheme uses (SBDF-q: q orders of
  the startup; RK: one per distinct H_ii; CN: the CNAB1 startup and the scheme's), each with its
  band rows, U, L, pivots, border V^T / W, capacitance S, their transposed twins, the solve
  work space of one right-hand side and, for a batch the library partitions (up to 256 systems
  long enough to split), the chunks' coupling spikes (csrc/banded.hip, create_impl and
  reserve_impl);
* explicit term: transform tables, the constant part of F and its workspace (RB 2D: 3 K Mz line
  products; channel: cn_ch3d_workspace_bytes).

Beyond the buffers, a ``runtime`` line of 10 % of the library's buffers (operators and factors)
covers what the device shows on top of them (allocation rounding, graph instantiations, kernel
scratch), as measured on MI355X: a 64 x 65 x 64 channel run under SBDF3 holds 2978 MiB where its
buffers are 2712 MiB, a RB 2D 256 x 128 run under SBDF2 234 MiB for 222 MiB
(tests/test_gpu_memory.py).  ``PROCESS_FLOOR_BYTES`` is what a process holds before any run (HIP
context, the code objects of libcnhip and torch, and the library's fixed 64 MiB work-space array
of the allocation-free solve, which every device holds once the library is loaded): up to 324 MiB
measured on MI355X, 352 MiB declared (the reference declares 512 MiB for its CUDA stack).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["MemoryPlan", "plan", "handle_bytes", "PROCESS_FLOOR_BYTES"]

MiB = 2 ** 20
# HIP context + code objects of a process that has loaded torch and libcnhip and run one solve,
# with the library's 64 MiB static work-space array: up to 324 MiB measured on MI355X (ROCm 7,
# torch 2.10; alone and after other runs), declared rounded up
PROCESS_FLOOR_BYTES = 352 * MiB

_KL_SET = (1, 2, 4, 6, 8, 10, 12, 14, 16)
_WU_SET = (4, 8, 16, 24, 32)


def _pick(v: int, options) -> int:
    for o in options:
        if o >= v:
            return o
    raise ValueError(f"band width {v} beyond the compiled kernels")


_LDS_CAP = 160 * 1024
_WAVE_SMALL = 16 + 8 * 8 * 16 + 8 * 4 + 16


def _a16(b: int) -> int:
    return (b + 15) // 16 * 16


def _wave_buf(KLP: int, US: int, esz: int, sp: int) -> int:
    return max(_a16(sp * KLP * esz + sp * 4), _a16(sp * US * esz))


def _parts(B: int, n: int, KLP: int, M: int, US: int, esz: int) -> tuple[int, int]:
    """(wavefronts per system, longest chunk) of the automatic kernel mode (csrc/banded.hip:
    wave mode when the one-wavefront LDS fits and the batch is not a lane-mode one, partitioned
    for up to 256 systems by pick_parts); (1, 0) when the batch is not partitioned."""
    lane_fast = KLP + 1 <= 8 and US - 1 <= 16
    wave_ok = _a16(n * esz) + _WAVE_SMALL + 3 * _wave_buf(KLP, US, esz, 32) <= _LDS_CAP
    if (lane_fast and B >= 1024) or not wave_ok or B > 256:
        return 1, 0
    wmax = max(KLP, M - 1)
    if wmax > 32:
        return 1, 0
    for P in (8, 4):
        if n // P < wmax:
            continue
        plen = (n + P - 1) // P
        for sp in (16, 8):
            nb = 2 if sp <= 8 else 3
            lds = (_a16(n * esz) + _WAVE_SMALL + P * (32 + 8) * esz
                   + P * (_a16((plen + wmax) * esz) + nb * _wave_buf(KLP, US, esz, sp)))
            if lds <= _LDS_CAP:
                return P, plen
    return 1, 0


def handle_bytes(B: int, n: int, kl: int, ku: int, r: int, esz: int = 16, nrhs: int = 1) -> int:
    """Device bytes of one factor handle of B systems in the automatic kernel mode (csrc/banded.hip
    create_impl: the kernels' padded widths KLP, WU, systems padded to whole 64-system workgroups,
    the coupling spikes of a partitioned batch; reserve_impl: the solve work space of ``nrhs``
    right-hand sides)."""
    KLP = _pick(max(kl, 1), _KL_SET)
    M = KLP + ku + 1
    US = _pick(KLP + ku, _WU_SET) + 1
    P, plen = _parts(B, n, KLP, M, US, esz)
    spikes = B * P * (KLP * (plen + KLP) + (M - 1) * (plen + M - 1)) * esz if P > 1 else 0
    Bp = (B + 63) // 64 * 64
    nV = max(Bp * r * n, 1)
    S = max(Bp * r * r, 1)
    sp_ = max(Bp * r, 1)
    total = (Bp * n * M * esz + Bp * n * US * esz + Bp * n * KLP * esz + Bp * n * 4      # Ab, Uf, Lf, piv
             + 2 * nV * esz + S * esz + sp_ * 4 + Bp * 8 + Bp * 4 + n * 4                 # Vt, W, S, spiv, cond, info, perm
             + nV * esz + S * esz + sp_ * 4 + Bp * 8 + Bp * 4)                            # transposed twins
    return total + spikes + B * nrhs * n * esz


def _implicit_operators(scheme: str, dt: float) -> int:
    """How many factor handles the scheme builds (distinct alpha M + beta L)."""
    from .runtime.imex import CN_SCHEMES, cn_coefficients, rk_tableau
    if scheme.startswith("SBDF"):
        return int(scheme[4])
    if scheme in CN_SCHEMES:
        keys = {cn_coefficients("CNAB1")[1][0] / cn_coefficients("CNAB1")[0][0]}
        a, b, _ = cn_coefficients(scheme)
        keys.add(b[0] / a[0])
        return len(keys)
    A, H, c = rk_tableau(scheme)
    return len({dt * H[i][i] for i in range(1, len(c))})


def _integrator_vectors(scheme: str, explicit: bool) -> int:
    """Device vectors of K n entries the integrator holds (imex.py, Imex.__init__)."""
    from .runtime.imex import CN_SCHEMES, cn_coefficients, rk_tableau
    if scheme.startswith("SBDF"):
        q = int(scheme[4])
        return 2 * q + 1 + (1 if explicit else 0)
    if scheme in CN_SCHEMES:
        return 3 * (len(cn_coefficients(scheme)[0]) - 1) + 1
    s = len(rk_tableau(scheme)[2]) - 1
    return 3 * s + 2


@dataclass
class MemoryPlan:
    lines: dict = field(default_factory=dict)        # category -> bytes

    @property
    def total(self) -> int:
        return int(sum(self.lines.values()))

    def tolerance_bytes(self) -> int:
        """What the plan may miss the device by: 64 MiB and 5 %."""
        return int(64 * MiB + 0.05 * self.total)

    def explain(self) -> str:
        rows = [f"  {k:<22s} {v / MiB:10.1f} MiB" for k, v in self.lines.items()]
        return "\n".join(["device memory plan"] + rows + [f"  {'total':<22s} {self.total / MiB:10.1f} MiB",
                                                           f"  {'process floor':<22s} {PROCESS_FLOOR_BYTES / MiB:10.1f} MiB"])


def plan(problem, scheme: str, dt: float = 1e-3, kl: int = 12, ku: int = 18) -> MemoryPlan:
    """The device bytes of one ``IVPSolver(problem, scheme, dt)`` after its first steps, by
    category (no GPU needed).  ``kl``, ``ku``: the band of the mode systems (12, 18 for the
    interleaved RB 2D and channel layouts)."""
    from .problems import Channel3D
    K, n, r = problem.K, problem.n, len(problem.border)
    es = 16
    vec = K * n * es
    p = MemoryPlan()
    p.lines["state"] = 2 * vec                                    # X and the continuation copy
    p.lines["integrator"] = _integrator_vectors(scheme, True) * vec
    nnz = sum(int(M.nnz) + int(L.nnz) for M, L in problem.mats)
    p.lines["operators (CSR)"] = nnz * (es + 4) + 2 * K * (n + 1) * 4
    p.lines["factors"] = _implicit_operators(scheme, dt) * handle_bytes(K, n, kl, ku, r, es)
    N = problem.N
    if isinstance(problem, Channel3D):
        tables = (problem.My * N * 2 * 8 + (problem.Mx + problem.Mz) * es + problem.My * N * 8 + 3 * N * 4)
    else:
        tables = 2 * N * problem.Mz * 8 + problem.Mx * es + problem.Mz * N * 8 + 3 * N * 4
    p.lines["explicit term"] = problem.explicit_scratch_bytes() + tables + vec
    p.lines["runtime"] = int(0.1 * (p.lines["operators (CSR)"] + p.lines["factors"]))
    return p
