"""Ensembles of 2D Rayleigh-Benard runs on the MI355X.

``Ensemble`` advances B members of one discretization together (the reference's Ensemble,
python/conatus/ensembles/ensemble.py:39-204, with BatchedSystem / BatchedRHS,
python/conatus/ensembles/batched.py:116-613): every member has the grid (Nx, Nz, Lx, walls,
dealiasing) of the first; Ra and Pr may differ per member (``from_solvers``).  A stage is one
explicit-term launch pair for all members (F does not depend on Ra or Pr), one fused SBDF/RK
right-hand-side kernel over the B x K stacked mode systems and one batched solve of all of them.
Members never interact: each member's arithmetic is the one of its single run, so a member is
bit-identical to the same run stepped alone by ``IVPSolver`` with the same kernel mode (tested).

``EnsembleSet`` advances runs of uneven sizes (several discretizations): one Ensemble per size
class, each stepped from its own captured HIP graphs on its own stream, so that the classes'
kernels overlap on the device.  The kernel mode of each class (lane mapping: one lane per system
and right-hand side with streamed sweeps; wave mapping: one wavefront per system and right-hand
side) is picked by timing both on the class's own factored systems (``autotune_mode``).
"""
from __future__ import annotations

import time
from collections.abc import Callable, Sequence

import numpy as np

from .errors import OptionError
from .problems import RayleighBenard2D
from .runtime.banded import SOLVE_MODES, BatchHandle, csr_args
from .runtime.imex import Imex, sbdf_coefficients
from .timestep import Channel3DExplicit, IVPSolver, RB2DExplicit, explicit_term

__all__ = ["Ensemble", "EnsembleSet", "autotune_mode", "discretization_key"]


def _torch():
    import torch
    return torch


def discretization_key(pb) -> tuple:
    """What members of one Ensemble must share: the grid and the dealiasing (not the parameters
    the implicit operators depend on: Ra, Pr of RB 2D, Re of the channel)."""
    return pb.discretization_key()


def _member_indices(member_of, nsolvers: int) -> np.ndarray:
    """The solver index of every member: whole numbers (bools are not indices) in [0, nsolvers)."""
    vals = list(member_of) if not isinstance(member_of, np.ndarray) else member_of.tolist()
    if isinstance(member_of, np.ndarray) and member_of.ndim != 1:
        vals = None
    ok = vals is not None and all(isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_))
                                  and 0 <= int(v) < nsolvers for v in vals)
    if not ok:
        raise OptionError(f"Ensemble.from_solvers: member_of={member_of!r} is not a list of solver indices "
                          f"(whole numbers from 0 to {nsolvers - 1})",
                          fix="member_of=[0, 0, 1, 1]: members 0 and 1 use solvers[0], 2 and 3 solvers[1]")
    return np.asarray([int(v) for v in vals], dtype=int)


class _LoopExplicit:
    """F member by member (the reference's rhs='loop' path): one single-member launch pair per
    member on its slice of the stacked state; the same kernels as the batched path."""

    def __init__(self, pb, members: int):
        self.one = explicit_term(pb)
        self.members = members
        self.size = pb.K * pb.n
        self._base = _torch().cat([self.one.base()] * members)

    def base(self):
        return self._base

    def write(self, x, out) -> None:
        xs, os_ = x.reshape(self.members, self.size), out.reshape(self.members, self.size)
        for b in range(self.members):
            self.one.write(xs[b], os_[b])


class Ensemble:
    """B members of one discretized 2D Rayleigh-Benard problem advanced together (module doc)."""

    def __init__(self, solver: IVPSolver, size: int, *, rhs: str = "batched"):
        if int(size) != size or size < 1:
            raise OptionError(f"Ensemble: size must be a whole number >= 1, got {size!r}")
        self._setup([solver], np.zeros(int(size), int), rhs)

    @classmethod
    def from_solvers(cls, solvers: Sequence[IVPSolver], *, rhs: str = "batched",
                     member_of: Sequence[int] | None = None) -> Ensemble:
        """Parameter ensemble: member b uses ``solvers[member_of[b]]`` (default: one member per
        solver).  The solvers must share the discretization, scheme, dt, pivoting and kernel mode."""
        self = cls.__new__(cls)
        mo = np.arange(len(solvers)) if member_of is None else _member_indices(member_of, len(solvers))
        self._setup(list(solvers), mo, rhs)
        return self

    @classmethod
    def from_problems(cls, problems: Sequence[RayleighBenard2D], scheme: str, dt: float, *,
                      rhs: str = "batched", pivot: bool = True, mode: str | None = None) -> Ensemble:
        """One member per problem (same discretization), without building single-run solvers."""
        self = cls.__new__(cls)
        self._setup_problems(list(problems), scheme, float(dt), rhs, pivot, mode)
        return self

    def _setup(self, solvers: list, member_of: np.ndarray, rhs: str) -> None:
        if len(member_of) < 1:
            raise OptionError("Ensemble: at least one member")
        s0 = solvers[0]
        for s in solvers[1:]:
            for a in ("scheme", "dt", "pivot", "mode"):
                if getattr(s, a) != getattr(s0, a):
                    raise OptionError(f"Ensemble.from_solvers: solvers differ in {a} "
                                      f"({getattr(s, a)!r} vs {getattr(s0, a)!r})",
                                      fix="build every member's solver with the same numerics")
        self.solvers, self.member_of = solvers, member_of
        self._setup_problems([solvers[i].problem for i in member_of], s0.scheme, s0.dt, rhs, s0.pivot, s0.mode)

    def _setup_problems(self, problems: list, scheme: str, dt: float, rhs: str, pivot: bool,
                        mode: str | None) -> None:
        torch = _torch()
        if rhs not in ("batched", "loop"):
            raise OptionError(f"Ensemble: rhs must be 'batched' or 'loop', got {rhs!r}",
                              fix="rhs='batched' (one launch pair for all members) or rhs='loop'")
        if not problems:
            raise OptionError("Ensemble: at least one member")
        p0 = problems[0]
        key = discretization_key(p0)
        for b, pb in enumerate(problems):
            if discretization_key(pb) != key:
                raise OptionError(f"Ensemble: member {b} has the discretization {discretization_key(pb)}, "
                                  f"member 0 has {key}",
                                  fix="one Ensemble per discretization; EnsembleSet groups uneven sizes")
        self.problems = problems
        self.B, self.K, self.n = len(problems), p0.K, p0.n
        self.scheme, self.dt, self.pivot, self.mode = scheme, float(dt), pivot, mode
        self.rhs_mode = rhs
        if rhs == "loop":
            self.explicit = _LoopExplicit(p0, self.B)
        elif isinstance(p0, RayleighBenard2D):
            self.explicit = RB2DExplicit(p0, self.B)
        else:
            self.explicit = Channel3DExplicit(p0, self.B)
        Ms = [M for pb in problems for (M, L) in pb.mats]
        Ls = [L for pb in problems for (M, L) in pb.mats]
        self.imex = Imex(Ms, Ls, p0.border, scheme, dt, pivot=pivot, dtype=np.complex128,
                         explicit=self.explicit, mode=mode)
        self.X = torch.zeros((self.B, self.K, self.n), dtype=torch.complex128, device="cuda")
        self.t = 0.0

    # ------------------------------------------------------------------ state
    @property
    def iteration(self) -> int:
        return self.imex.iteration

    def _flat(self):
        return self.X.view(self.B * self.K, self.n)

    def set_states(self, X, t: float = 0.0) -> None:
        """Initial states (B, K, n) (the (K, n) mode states of the single runs); resets the
        multistep history."""
        X = np.asarray(X, dtype=np.complex128)
        if X.shape != (self.B, self.K, self.n):
            raise OptionError(f"Ensemble.set_states: expected shape {(self.B, self.K, self.n)}, got {X.shape}")
        self.X.copy_(_torch().from_numpy(np.ascontiguousarray(X)))
        self.t = float(t)
        self.imex.reset()

    def states(self) -> np.ndarray:
        """Host copy of the member states (B, K, n)."""
        return self.X.cpu().numpy()

    # ------------------------------------------------------------------ steps
    def step(self) -> None:
        self.t = self.imex.run(self._flat(), self.t, 1)

    def run(self, steps: int, callback: Callable[[Ensemble], None] | None = None,
            every: int | None = None) -> None:
        """Advance ``steps`` steps; ``callback(self)`` after every step whose iteration is a
        multiple of ``every`` (both explicit, as in the reference's Ensemble.run)."""
        if (callback is None) != (every is None):
            raise OptionError("Ensemble.run: callback and every go together")
        if every is not None and (int(every) != every or every < 1):
            raise OptionError("Ensemble.run: every must be a whole number >= 1")
        if callback is None:
            self.t = self.imex.run(self._flat(), self.t, steps)
            return
        for _ in range(int(steps)):
            self.step()
            if self.iteration % every == 0:
                callback(self)

    def run_graph(self, steps: int) -> float:
        """As ``run``, full-order steps replayed from captured HIP graphs."""
        self.t = self.imex.run_graph(self._flat(), self.t, steps)
        return self.t

    def explain(self) -> str:
        pb = self.problems[0]
        ps = sorted({p.params() for p in self.problems})
        return (f"Ensemble({self.B} member(s) of {pb.describe()}, parameters {ps[0]}..{ps[-1]}, "
                f"{self.scheme}, dt={self.dt:g}, rhs={self.rhs_mode}): {self.B * self.K} mode systems of "
                f"n={self.n} in one batched solve per stage; F in "
                + ("one launch sequence for all members" if self.rhs_mode == "batched"
                   else "one launch sequence per member")
                + "; " + self.imex.explain())


def autotune_mode(problems: Sequence[RayleighBenard2D], scheme: str, dt: float, reps: int = 5,
                  pivot: bool = True) -> tuple[str, dict]:
    """Kernel mode for one size class: the full-order operator a0/dt M + L of every member's modes
    factored in lane mode, wave mode with staged blocks of 8 and of 32 steps, and partitioned wave
    mode, ``reps`` timed solves of each; returns the fastest mode
    and the timings (ms per solve).  A mode that does not fit the shape is refused by the library
    and skipped here."""
    torch = _torch()
    q = {"SBDF1": 1, "SBDF2": 2, "SBDF3": 3, "SBDF4": 4}.get(scheme)
    alpha = sbdf_coefficients(q)[0][0] / dt if q else 1.0
    mats = [csr_args((alpha * M + L).tocsr(), np.complex128) for pb in problems for (M, L) in pb.mats]
    border = problems[0].border
    n = problems[0].n
    times: dict = {}
    rng = np.random.default_rng(0)
    b = torch.from_numpy(rng.standard_normal((len(mats), n)) + 1j * rng.standard_normal((len(mats), n))).cuda()
    for mode in ("lane", "wave8", "wave32", "part"):
        h = BatchHandle(mats, border, np.complex128, pivot, mode=mode)
        got = SOLVE_MODES[h.props().solve_mode]
        if got != mode.rstrip("0123456789"):
            continue
        h.reserve(1)
        d = b.clone()
        s = torch.cuda.current_stream().cuda_stream
        for _ in range(2):
            h.solve_ptr(d.data_ptr(), 1, d.numel(), False, s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            h.solve_ptr(d.data_ptr(), 1, d.numel(), False, s)
        torch.cuda.synchronize()
        times[mode] = (time.perf_counter() - t0) * 1e3 / reps
        del h
    if not times:
        raise OptionError("autotune_mode: no kernel mode fits these systems")
    return min(times, key=times.get), times


class EnsembleSet:
    """Runs of uneven sizes advanced together: one Ensemble per discretization (size class), each
    class stepped from its own captured HIP graphs on its own stream (module doc), so that the
    classes' kernels overlap on the device: a small class's launches fill the gaps the large
    classes leave instead of waiting for them."""

    def __init__(self, problems: Sequence[RayleighBenard2D], scheme: str, dt: float, *,
                 autotune: bool = True, mode: str | None = None, pivot: bool = True):
        if not problems:
            raise OptionError("EnsembleSet: at least one run")
        if autotune and mode is not None:
            raise OptionError("EnsembleSet: autotune and a fixed mode exclude each other")
        self.scheme, self.dt = scheme, float(dt)
        classes: dict = {}
        for i, pb in enumerate(problems):
            classes.setdefault(discretization_key(pb), []).append(i)
        self.groups: list[Ensemble] = []
        self.tuning: list[dict] = []
        self.where: list[tuple[int, int]] = [(0, 0)] * len(problems)     # run -> (group, member)
        for key in sorted(classes):
            idx = classes[key]
            probs = [problems[i] for i in idx]
            m, times = autotune_mode(probs, scheme, dt, pivot=pivot) if autotune else (mode, {})
            g = Ensemble.from_problems(probs, scheme, dt, mode=m, pivot=pivot)
            for b, i in enumerate(idx):
                self.where[i] = (len(self.groups), b)
            self.groups.append(g)
            self.tuning.append({"class": f"{key[0]}x{key[1]}", "runs": len(idx), "mode": m, "ms_per_solve": times})
        self.runs = len(problems)
        self.t = 0.0
        self.iteration = 0
        self._streams: list | None = None

    def set_states(self, states: Sequence) -> None:
        """Initial (K, n) state of every run, in the order of the problems; resets every history."""
        if len(states) != self.runs:
            raise OptionError(f"EnsembleSet.set_states: {self.runs} states expected, got {len(states)}")
        for gi, g in enumerate(self.groups):
            X = np.stack([np.asarray(states[i]) for i in range(self.runs) if self.where[i][0] == gi])
            g.set_states(X)
        self.t, self.iteration = 0.0, 0

    def states(self) -> list:
        per = [g.states() for g in self.groups]
        return [per[gi][b] for gi, b in self.where]

    def run(self, steps: int) -> None:
        """Eager steps, class after class."""
        for g in self.groups:
            g.run(steps)
        self.t = self.groups[0].t
        self.iteration += int(steps)

    def run_graph(self, steps: int) -> None:
        """Every class runs ``steps`` steps (lower-order startup eagerly, then full-order steps
        replayed from its captured graphs) on its own stream, ordered after the work already on
        the current stream; the current stream then waits for all of them.  The classes share
        nothing, so each one's arithmetic is that of its own Ensemble.run_graph."""
        torch = _torch()
        cur = torch.cuda.current_stream()
        if self._streams is None:
            self._streams = [torch.cuda.Stream() for _ in self.groups]
        for g, sc in zip(self.groups, self._streams):
            sc.wait_stream(cur)
            with torch.cuda.stream(sc):
                g.run_graph(steps)
        for sc in self._streams:
            cur.wait_stream(sc)
        self.t = self.groups[0].t
        self.iteration += int(steps)

    def explain(self) -> str:
        lines = [f"EnsembleSet({self.runs} runs in {len(self.groups)} size classes, {self.scheme}, dt={self.dt:g}): "
                 "each class from its own captured graphs on its own stream"]
        for g, tu in zip(self.groups, self.tuning):
            lines.append(f"  class {tu['class']}: {tu['runs']} runs, mode {tu['mode'] or 'automatic'} "
                         + (f"(ms per solve: {', '.join(f'{k} {v:.3f}' for k, v in tu['ms_per_solve'].items())})"
                            if tu["ms_per_solve"] else ""))
        return "\n".join(lines)
