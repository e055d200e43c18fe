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
from .problems import NBORDER, Channel3D, RayleighBenard2D
from .runtime.imex import Imex

__all__ = ["RB2DExplicit", "Channel3DExplicit", "IVPSolver"]


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
        # scratch: 3 K Mz line products and 6 K Mz line values per member
        self.Gh = torch.empty(self.members * 9 * K * Mz, dtype=torch.complex128, device=dev)
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


class Channel3DExplicit:
    """F(X) of the 3D plane channel on the device (cn_ch3d_explicit: five transform launches,
    one workspace).  The constant part (the driving pressure gradient) is added by the last
    launch; the other rows of F keep base()."""

    def __init__(self, pb: Channel3D):
        torch = _torch()
        self.pb = pb
        N, Mx, My, Mz = pb.N, pb.Mx, pb.My, pb.Mz
        dev = "cuda"
        By = np.stack([S.cheb_backward_matrix(N, My, 0), S.cheb_backward_matrix(N, My, 1, 2.0)], axis=-1)
        self.By = torch.from_numpy(np.ascontiguousarray(By)).to(dev)                       # (My, N, 2)
        self.Wx = torch.from_numpy(np.exp(2j * np.pi * np.arange(Mx) / Mx)).to(dev)
        self.Wz = torch.from_numpy(np.exp(2j * np.pi * np.arange(Mz) / Mz)).to(dev)
        Ay = (pb.S02 @ S.cheb_forward_matrix(N, My)).T
        self.Ay = torch.from_numpy(np.ascontiguousarray(Ay)).to(dev)                       # (My, N)
        self._base = torch.from_numpy(pb.b0().ravel()).to(dev)
        fidx = np.stack([pb.field_index(f) for f in (0, 1, 2)]).astype(np.int32)
        self.fidx = torch.from_numpy(np.ascontiguousarray(fidx)).to(dev)
        nbytes = _lib.lib().cn_ch3d_workspace_bytes(pb.Nx, pb.Ny, pb.Nz, Mx, My, Mz)
        self.ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)

    def base(self):
        return self._base

    def write(self, x, out) -> None:
        pb = self.pb
        v = ctypes.c_void_p
        p = lambda t: v(t.data_ptr())   # noqa: E731
        st = _lib.lib().cn_ch3d_explicit(p(x), p(self.fidx), p(self._base), p(out), p(self.ws), self.ws.numel(),
                                         p(self.By), p(self.Wx), p(self.Wz), p(self.Ay), pb.Nx, pb.Ny, pb.Nz,
                                         pb.Mx, pb.My, pb.Mz, pb.n, NBORDER, float(pb.kx0), float(pb.kz0),
                                         v(_torch().cuda.current_stream().cuda_stream))
        _lib.check(st, "Channel3DExplicit")


def explicit_term(problem):
    """The device explicit term of a problem."""
    if isinstance(problem, Channel3D):
        return Channel3DExplicit(problem)
    if isinstance(problem, RayleighBenard2D):
        return RB2DExplicit(problem)
    raise TypeError(f"no device explicit term for {type(problem).__name__}")


class IVPSolver:
    """Device IVP solver of a Fourier x Chebyshev problem (2D Rayleigh-Benard, 3D channel)."""

    def __init__(self, problem, scheme: str, dt: float, pivot: bool = True, mode: str | None = None):
        torch = _torch()
        self.problem = problem
        self.explicit = explicit_term(problem)
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
        if isinstance(pb, Channel3D):
            return (f"IVPSolver(channel 3D {pb.Nx}x{pb.Ny}x{pb.Nz}, Re={pb.Re:g}, {self.scheme}, dt={self.dt:g}): "
                    f"{pb.K} complex modes of n={pb.n} in one batched solve; F on the {pb.Mx}x{pb.My}x{pb.Mz} grid; "
                    + self.imex.explain())
        return (f"IVPSolver(RB 2D {pb.Nx}x{pb.Nz}, Ra={pb.Ra:g}, Pr={pb.Pr:g}, {self.scheme}, dt={self.dt:g}): "
                f"{pb.K} complex modes of n={pb.n} in one batched solve; F on the {pb.Mx}x{pb.Mz} grid; "
                + self.imex.explain())
