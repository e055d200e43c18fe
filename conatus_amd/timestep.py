"""IVP time stepping of the Fourier x Chebyshev problems on the MI355X.

``IVPSolver(problem, scheme, dt)`` mirrors the reference's ``IVPSolver`` step loop
(python/conatus/timestep/integrator.py:648-682, 361-373, 608-646): the stacked state of all Fourier
modes stays resident in HBM; a step is the explicit term F(X) (the fused transform/product
pipeline of libcnhip), the fused SBDF / RK right-hand-side assembly, and one batched
bordered-banded solve of all modes; ``run_graph`` replays full-order steps from captured HIP
graphs (integrator.py:781-954).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import spectral as S
from .problems import NBORDER, RayleighBenard2D
from .runtime.imex import Imex

__all__ = ["RB2DExplicit", "IVPSolver"]


def _torch():
    import torch
    return torch


class RB2DExplicit:
    """F(X) of 2D Rayleigh-Benard on the device (cn_rb2d_explicit_many): `members` states of one
    discretization stored one after another (members x K x n), one launch pair for all of them.
    F does not depend on Ra or Pr, so members with different parameters share it."""

    def __init__(self, pb: RayleighBenard2D, members: int = 1):
        torch = _torch()
        self.pb = pb
        self.members = int(members)
        K, N, Mx, Mz = pb.K, pb.N, pb.Mx, pb.Mz
        dev = "cuda"
        Bz = np.stack([S.cheb_backward_matrix(N, Mz, 0, pb.Lz).T, S.cheb_backward_matrix(N, Mz, 1, pb.Lz).T])
        self.Bz = torch.from_numpy(np.ascontiguousarray(Bz)).to(dev)                       # (2, N, Mz)
        self.Wx = torch.from_numpy(np.exp(2j * np.pi * np.arange(Mx) / Mx)).to(dev)        # twiddles
        Az = (pb.S02 @ S.cheb_forward_matrix(N, Mz)).T
        self.Az = torch.from_numpy(np.ascontiguousarray(Az)).to(dev)                       # (Mz, N)
        self.Gh = torch.empty(self.members * 3 * K * Mz, dtype=torch.complex128, device=dev)   # scratch
        self._base = torch.from_numpy(np.tile(pb.b0().ravel(), self.members)).to(dev)
        fidx = np.stack([pb.field_index(f) for f in (0, 1, 2)]).astype(np.int32)
        self.fidx = torch.from_numpy(np.ascontiguousarray(fidx)).to(dev)

    def base(self):
        return self._base

    def write(self, x, out) -> None:
        pb = self.pb
        v = ctypes.c_void_p
        p = lambda t: v(t.data_ptr())   # noqa: E731
        st = _lib.lib().cn_rb2d_explicit_many(p(x), p(self.fidx), p(out), p(self.Gh), p(self.Bz), p(self.Wx),
                                              p(self.Az), pb.K, pb.N, pb.Mx, pb.Mz, pb.n, NBORDER, float(pb.kappa),
                                              self.members, v(_torch().cuda.current_stream().cuda_stream))
        _lib.check(st, "RB2DExplicit")


class IVPSolver:
    """Device IVP solver of a Fourier x Chebyshev problem (2D Rayleigh-Benard)."""

    def __init__(self, problem: RayleighBenard2D, scheme: str, dt: float, pivot: bool = True,
                 mode: str | None = None):
        torch = _torch()
        self.problem = problem
        self.explicit = RB2DExplicit(problem)
        Ms = [m[0] for m in problem.mats]
        Ls = [m[1] for m in problem.mats]
        self.imex = Imex(Ms, Ls, problem.border, scheme, dt, pivot=pivot, dtype=np.complex128,
                         explicit=self.explicit, mode=mode)
        self.pivot, self.mode = pivot, mode
        self.X = torch.zeros((problem.K, problem.n), dtype=torch.complex128, device="cuda")
        self.t = 0.0
        self.scheme, self.dt = scheme, float(dt)

    @property
    def iteration(self) -> int:
        return self.imex.iteration

    def set_state(self, X: np.ndarray, t: float = 0.0) -> None:
        self.X.copy_(_torch().from_numpy(np.ascontiguousarray(X, dtype=np.complex128)))
        self.t = float(t)
        self.imex.reset()

    def state(self) -> np.ndarray:
        return self.X.cpu().numpy()

    def step(self) -> None:
        self.t = self.imex.run(self.X, self.t, 1)

    def run(self, steps: int) -> float:
        self.t = self.imex.run(self.X, self.t, steps)
        return self.t

    def run_graph(self, steps: int) -> float:
        self.t = self.imex.run_graph(self.X, self.t, steps)
        return self.t

    def save_checkpoint(self, path):
        """State + SBDF history + clock (``conatus.checkpoint/1``, ``.npz``): a restart with
        ``load_checkpoint`` continues bit for bit (integrator.py:997-1021)."""
        from .checkpoint import save_checkpoint
        return save_checkpoint(self, path)

    def load_checkpoint(self, path) -> None:
        """Continue from ``save_checkpoint``; the numerics and the problem must match
        (integrator.py:1023-1076)."""
        from .checkpoint import load_checkpoint
        load_checkpoint(self, path)

    def explain(self) -> str:
        pb = self.problem
        return (f"IVPSolver(RB 2D {pb.Nx}x{pb.Nz}, Ra={pb.Ra:g}, Pr={pb.Pr:g}, {self.scheme}, dt={self.dt:g}): "
                f"{pb.K} complex modes of n={pb.n} in one batched solve; F on the {pb.Mx}x{pb.Mz} grid; "
                + self.imex.explain())
