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
from .errors import NumericalError, OptionError
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
    """F(X) of the 3D plane channel on the device (cn_ch3d_explicit_many: five transform launches
    for all `members` states of one grid stored one after another, a workspace region per member).
    The constant part (the driving pressure gradient) is added by the last launch; the other rows
    of F keep base()."""

    def __init__(self, pb: Channel3D, members: int = 1):
        torch = _torch()
        self.pb = pb
        self.members = int(members)
        N, Mx, My, Mz = pb.N, pb.Mx, pb.My, pb.Mz
        dev = "cuda"
        By = np.stack([S.cheb_backward_matrix(N, My, 0), S.cheb_backward_matrix(N, My, 1, 2.0)], axis=-1)
        self.By = torch.from_numpy(np.ascontiguousarray(By)).to(dev)                       # (My, N, 2)
        self.Wx = torch.from_numpy(np.exp(2j * np.pi * np.arange(Mx) / Mx)).to(dev)
        self.Wz = torch.from_numpy(np.exp(2j * np.pi * np.arange(Mz) / Mz)).to(dev)
        Ay = (pb.S02 @ S.cheb_forward_matrix(N, My)).T
        self.Ay = torch.from_numpy(np.ascontiguousarray(Ay)).to(dev)                       # (My, N)
        self._one_base = torch.from_numpy(pb.b0().ravel()).to(dev)
        self._base = self._one_base if self.members == 1 else torch.cat([self._one_base] * self.members)
        fidx = np.stack([pb.field_index(f) for f in (0, 1, 2)]).astype(np.int32)
        self.fidx = torch.from_numpy(np.ascontiguousarray(fidx)).to(dev)
        nbytes = _lib.lib().cn_ch3d_workspace_bytes(pb.Nx, pb.Ny, pb.Nz, Mx, My, Mz)
        self.ws = torch.empty(max(int(nbytes) * self.members, 1), dtype=torch.uint8, device=dev)

    def base(self):
        return self._base

    def write(self, x, out) -> None:
        pb = self.pb
        v = ctypes.c_void_p
        p = lambda t: v(t.data_ptr())   # noqa: E731
        st = _lib.lib().cn_ch3d_explicit_many(p(x), p(self.fidx), p(self._one_base), p(out), p(self.ws),
                                              self.ws.numel(), p(self.By), p(self.Wx), p(self.Wz), p(self.Ay), pb.Nx,
                                              pb.Ny, pb.Nz, pb.Mx, pb.My, pb.Mz, pb.n, NBORDER, float(pb.kx0),
                                              float(pb.kz0), self.members,
                                              v(_torch().cuda.current_stream().cuda_stream))
        _lib.check(st, "Channel3DExplicit")


def explicit_term(problem):
    """The device explicit term of a problem."""
    if isinstance(problem, Channel3D):
        return Channel3DExplicit(problem)
    if isinstance(problem, RayleighBenard2D):
        return RB2DExplicit(problem)
    raise TypeError(f"no device explicit term for {type(problem).__name__}")


#: Steps between two finite-state checks of an IVP run (the reference's FINITE_CHECK_EVERY,
#: python/conatus/timestep/run_checks.py:14-20): one device reduction and host synchronisation.
FINITE_CHECK_EVERY = 100


def finite_cadence(check_finite, where: str) -> int:
    """True -> every FINITE_CHECK_EVERY steps, False -> off (0), a positive int -> that cadence."""
    if check_finite is True:
        return FINITE_CHECK_EVERY
    if check_finite is False:
        return 0
    if isinstance(check_finite, (int, np.integer)) and not isinstance(check_finite, bool) and check_finite > 0:
        return int(check_finite)
    raise OptionError(f"{where}: check_finite={check_finite!r}",
                      why=f"the check is on (True: every {FINITE_CHECK_EVERY} steps), off (False) or "
                          "every n steps (n > 0)",
                      fix="check_finite=True (default), check_finite=10 or check_finite=False")


class IVPSolver:
    """Device IVP solver of a Fourier x Chebyshev problem (2D Rayleigh-Benard, 3D channel).

    The clock follows the reference's rule t = t_origin + (iteration - it_origin) dt
    (integrator.py:680): the origin is the (t, iteration) of ``set_state`` or of a restart, so no
    round-off accumulates over a long run.  ``check_finite`` (default every 100 steps) tests the
    state for NaN/inf at multiples of the cadence and at the end of every ``run``/``run_graph``
    call and raises ``NumericalError`` naming the iteration, the time and the unknown."""

    def __init__(self, problem, scheme: str, dt: float, pivot: bool = True, mode: str | None = None,
                 check_finite=True):
        torch = _torch()
        self.problem = problem
        self._finite_every = finite_cadence(check_finite, f"IVPSolver(problem, {scheme!r}, ...)")
        self.explicit = explicit_term(problem)
        Ms = [m[0] for m in problem.mats]
        Ls = [m[1] for m in problem.mats]
        self.imex = Imex(Ms, Ls, problem.border, scheme, dt, pivot=pivot, dtype=np.complex128,
                         explicit=self.explicit, mode=mode)
        self.pivot, self.mode = pivot, mode
        self.X = torch.zeros((problem.K, problem.n), dtype=torch.complex128, device="cuda")
        self.scheme, self.dt = scheme, float(dt)
        self.t = 0.0
        self.t_origin, self.it_origin = 0.0, 0
        self._finite_ok = None

    @property
    def iteration(self) -> int:
        return self.imex.iteration

    def _clock(self, iteration: int) -> float:
        return self.t_origin + (iteration - self.it_origin) * self.dt

    def set_state(self, X: np.ndarray, t: float = 0.0) -> None:
        """Start a new trajectory from the (K, n) state X at time t (history and iteration count
        reset; the time origin becomes (t, 0))."""
        self.X.copy_(_torch().from_numpy(np.ascontiguousarray(X, dtype=np.complex128)))
        self.imex.reset()
        self.t = float(t)
        self.t_origin, self.it_origin = float(t), 0
        self._finite_ok = None

    def state(self) -> np.ndarray:
        return self.X.cpu().numpy()

    def _steps_arg(self, steps, where: str) -> int:
        if isinstance(steps, bool) or not isinstance(steps, (int, np.integer)) or steps < 0:
            raise OptionError(f"{where}: steps={steps!r} is not a number of steps",
                              fix="an integer >= 0, e.g. steps=100")
        return int(steps)

    def step(self) -> None:
        self.t = self.imex.run(self.X, self.t, 1, self._clock)
        if self._finite_every and self.iteration % self._finite_every == 0:
            self.check_finite("IVPSolver.step")

    def run(self, steps: int) -> float:
        """Advance ``steps`` steps; the state is checked finite at every multiple of the cadence
        and at the end of the call."""
        steps = self._steps_arg(steps, "IVPSolver.run")
        fe, left = self._finite_every, steps
        while left:
            k = min(left, fe - self.iteration % fe) if fe else left
            self.t = self.imex.run(self.X, self.t, k, self._clock)
            left -= k
            if fe and self.iteration % fe == 0 and left:
                self.check_finite("IVPSolver.run")
        if fe and steps:
            self.check_finite("IVPSolver.run (end of the call)")
        return self.t

    def run_graph(self, steps: int) -> float:
        """As ``run``, full-order steps replayed from captured HIP graphs; one finite-state check
        at the end of the call (no host synchronisation between replays)."""
        steps = self._steps_arg(steps, "IVPSolver.run_graph")
        self.t = self.imex.run_graph(self.X, self.t, steps, clock=self._clock)
        if self._finite_every and steps:
            self.check_finite("IVPSolver.run_graph (end of the call)")
        return self.t

    def check_finite(self, where: str = "IVPSolver.check_finite") -> None:
        """Raise NumericalError if the state holds a NaN or an inf (one device reduction):
        names the iteration, the time, the first non-finite unknown and the last iteration at
        which the state was finite (python/conatus/timestep/run_checks.py:64-97)."""
        torch = _torch()
        if bool(torch.isfinite(self.X).all()):
            self._finite_ok = (self.iteration, self.t)
            return
        bad = ~torch.isfinite(self.X).cpu().numpy()
        name, count = self.problem.unknown_name(int(np.flatnonzero(bad.any(axis=0))[0])), int(bad.sum())
        last = self._finite_ok
        since = f"; last checked finite at iteration {last[0]} (t = {last[1]:.10g})" if last else ""
        raise NumericalError(
            f"non-finite state (NaN or inf) at iteration {self.iteration}, t = {self.t:.10g}: "
            f"unknown {name!r} ({count} non-finite coefficient(s)){since}", where=where,
            why="the run has blown up (a numerical or physical instability the resolution cannot follow)",
            fix=f"reduce dt (now {self.dt:g}), raise the resolution, or check the initial state; "
                "check_finite=False turns this check off")

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
