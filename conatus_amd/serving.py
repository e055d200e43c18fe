"""Serving of IVP runs on one MI355X: continuous batching of 2D Rayleigh-Benard requests.

Requests (a problem, an initial state, a number of steps) arrive at any time.  The server keeps
every admitted run's state, SBDF history and factored mode systems resident in HBM and advances
all of them in ticks; one tick is one step of every active run.

* Size classes: every discretization (grid) has one resident slot ensemble, an ``Ensemble`` of
  ``capacity`` slots whose full-order steps are replayed from its captured graphs (one batched
  explicit term, stage and solve for all slots).  The kernel mode of a class is picked once by
  ``autotune_mode``.
* Admission, at the start of a tick: waiting requests are admitted in arrival order while the
  estimated HBM footprint of the admitted runs stays within the budget.  The runs admitted in one
  tick make their lower-order startup steps together in a startup cohort (one Ensemble per class)
  on its own stream, beside the slot ensembles.
* Joining: a run whose startup is complete moves into a free slot of its class by device copies
  only: its state, its SBDF history levels (into the ring slots the class's next step reads) and
  its factored full-order mode systems (``cn_bb_copy_systems``; skipped when the slot already
  holds the factors of the same Ra and Pr).  From then on it shares the class's batched step.
  A full class grows to twice the slots (states, history and factors copied over).
* Retirement: a run that reached its step count leaves with its state copied to the host; its
  slot is free for the next run (a free slot keeps being stepped, its values are never read).

Moving a run only copies values: every run's arithmetic stays the one of its single run (tested:
each result is the same alone and in the mix).  The step latency of a run is the wall time of
the tick it advanced in (admission, joins, every class's step and every startup step, until the
device finished).  There is no counterpart in the reference (it has no server); the per-run step
is the reference's IVPSolver step (python/conatus/timestep/integrator.py:648-682).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .ensembles import Ensemble, autotune_mode, discretization_key
from .errors import OptionError
from .problems import RayleighBenard2D
from .runtime.imex import sbdf_coefficients

__all__ = ["Request", "Server", "footprint_bytes"]


def _torch():
    import torch
    return torch


@dataclass
class Request:
    problem: RayleighBenard2D
    steps: int
    X0: np.ndarray
    arrival: float = 0.0                  # seconds (or ticks) after the start of the serving loop
    rid: int = -1
    # filled by the server
    admitted: float | None = None
    finished: float | None = None
    result: np.ndarray | None = field(default=None, repr=False)
    step_latencies: list = field(default_factory=list, repr=False)


def footprint_bytes(pb: RayleighBenard2D, order: int = 2) -> int:
    """Estimated HBM bytes of one resident run: state, SBDF history (M X and F per level), the
    explicit-term scratch and the factor storage of its K mode systems (band rows, U, L, W, V^T,
    pivots; the largest band a RB 2D mode system has: kl = 12, ku = 18, r = 7)."""
    K, n = pb.K, pb.n
    vec = K * n * 16
    kl, ku, r = 12, 18, 7
    factor = K * n * 16 * ((kl + ku + 1) + (kl + ku + 1) + kl + 2 * r) + K * n * 4
    return vec * (1 + 2 * order + 3) + factor


def _params(pb: RayleighBenard2D) -> tuple:
    return (float(pb.Ra), float(pb.Pr))


def _full_order_factor(imex):
    """The factored full-order operator a0/dt M + L of an SBDF integrator."""
    return imex._factor(sbdf_coefficients(imex.order)[0][0] / imex.dt, 1.0)


class _SlotClass:
    """The resident slot ensemble of one size class."""

    def __init__(self, template: RayleighBenard2D, capacity: int, scheme: str, dt: float, mode: str | None):
        torch = _torch()
        self.template, self.scheme, self.dt, self.mode = template, scheme, dt, mode
        self.stream = torch.cuda.Stream()
        self.members: dict = {}
        self.slot_params: list = []
        self.capacity = 0
        self.ens = None
        self.fac = None
        self._grow(capacity)

    def _grow(self, capacity: int) -> None:
        """A slot ensemble of ``capacity`` slots (the template's problem in every slot, full order
        from the start: runs join with their own history); the current slots' states, history
        levels and factors are carried over into its first slots."""
        torch = _torch()
        old_ens, old_fac, oc = self.ens, self.fac, self.capacity
        ens = Ensemble.from_problems([self.template] * capacity, self.scheme, self.dt, mode=self.mode)
        im = ens.imex
        K, n = ens.K, ens.n
        with torch.cuda.stream(self.stream):
            fac = _full_order_factor(im)
            im._levels = list(range(im.order))
            if old_ens is not None:
                oim = old_ens.imex
                ens._flat()[: oc * K].copy_(old_ens._flat())
                for i, s_old in enumerate(oim._levels):
                    for j in range(2):
                        im._ring[i][j].view(im.batch, n)[: oc * K].copy_(oim._ring[s_old][j].view(oim.batch, n))
                fac.copy_systems_from(old_fac, 0, 0, oc * K, self.stream.cuda_stream)
            ens.t = 0.0
        self.slot_params = self.slot_params + [_params(self.template)] * (capacity - oc)
        self.ens, self.fac, self.capacity = ens, fac, capacity

    def free_slot(self) -> int:
        for s in range(self.capacity):
            if s not in self.members:
                return s
        self._grow(2 * self.capacity)
        return self.free_slot()

    def join(self, cohort: Ensemble, b: int, req: Request, cohort_fac) -> None:
        """Move member b of a startup cohort (history complete) into a free slot."""
        torch = _torch()
        s = self.free_slot()
        ens, im = self.ens, self.ens.imex
        K, n = ens.K, ens.n
        cim = cohort.imex
        with torch.cuda.stream(self.stream):
            self.stream.wait_stream(torch.cuda.current_stream())
            ens._flat()[s * K:(s + 1) * K].copy_(cohort._flat()[b * K:(b + 1) * K])
            # history level i of the run goes to the ring slot that is level i of the class now:
            # the class's next step reads it as level i + 1
            for i, cs in enumerate(cim._levels):
                for j in range(2):
                    im._ring[im._levels[i]][j].view(im.batch, n)[s * K:(s + 1) * K].copy_(
                        cim._ring[cs][j].view(cim.batch, n)[b * K:(b + 1) * K])
            if self.slot_params[s] != _params(req.problem):
                self.fac.copy_systems_from(cohort_fac, b * K, s * K, K, self.stream.cuda_stream)
                self.slot_params[s] = _params(req.problem)
        self.members[s] = req

    def step(self) -> None:
        torch = _torch()
        with torch.cuda.stream(self.stream):
            self.ens.imex.replay_step(self.ens._flat())
        self.ens.t += self.ens.dt


class Server:
    """Continuous-batching server of 2D Rayleigh-Benard IVP requests (module doc)."""

    def __init__(self, scheme: str = "SBDF2", dt: float = 1e-4, hbm_budget: float = 200e9,
                 autotune: bool = True, mode: str | None = None, capacity: int = 8):
        if autotune and mode is not None:
            raise OptionError("Server: autotune and a fixed mode exclude each other")
        if scheme not in ("SBDF1", "SBDF2", "SBDF3", "SBDF4"):
            raise OptionError(f"Server: scheme {scheme!r} not served", fix="one of SBDF1..SBDF4")
        if int(capacity) != capacity or capacity < 1:
            raise OptionError("Server: capacity must be a whole number >= 1")
        self.scheme, self.dt = scheme, float(dt)
        self.budget = float(hbm_budget)
        self.autotune, self.mode = autotune, mode
        self.capacity = int(capacity)
        self._order = int(scheme[-1])
        self._modes: dict = {}
        self.classes: dict = {}
        self.waiting: list[Request] = []
        self.startups: list = []          # (Ensemble, [requests], stream)
        self.finished: list[Request] = []
        self.resident = 0
        self.ticks: list[float] = []
        self.joins = 0
        self._next_id = 0

    # ------------------------------------------------------------------ requests
    def submit(self, req: Request) -> Request:
        if int(req.steps) != req.steps or req.steps < 1:
            raise OptionError("Server.submit: steps must be a whole number >= 1")
        X0 = np.asarray(req.X0)
        if X0.shape != (req.problem.K, req.problem.n):
            raise OptionError(f"Server.submit: X0 must have shape {(req.problem.K, req.problem.n)}")
        if footprint_bytes(req.problem, self._order) > self.budget:
            raise OptionError("Server.submit: the run alone exceeds the HBM budget")
        req.rid = self._next_id
        self._next_id += 1
        self.waiting.append(req)
        return req

    def _mode_for(self, pb: RayleighBenard2D) -> str | None:
        if not self.autotune:
            return self.mode
        key = discretization_key(pb)
        if key not in self._modes:
            self._modes[key] = autotune_mode([pb], self.scheme, self.dt)[0]
        return self._modes[key]

    def _class(self, pb: RayleighBenard2D) -> _SlotClass:
        key = discretization_key(pb)
        c = self.classes.get(key)
        if c is None:
            c = _SlotClass(pb, self.capacity, self.scheme, self.dt, self._mode_for(pb))
            self.classes[key] = c
        return c

    def prepare(self, problems) -> None:
        """Tune and build the size classes of ``problems`` ahead of the requests."""
        for pb in problems:
            self._class(pb)

    # ------------------------------------------------------------------ scheduling
    def _admit(self, now: float) -> None:
        torch = _torch()
        take: list[Request] = []
        rest: list[Request] = []
        blocked = False
        for q in self.waiting:
            if q.arrival > now:
                rest.append(q)
                continue
            fp = footprint_bytes(q.problem, self._order)
            if not blocked and self.resident + fp <= self.budget:
                self.resident += fp
                q.admitted = now
                take.append(q)
            else:
                blocked = True          # arrival order: nothing overtakes a request that waits
                rest.append(q)
        self.waiting = rest
        by: dict = {}
        for q in take:
            by.setdefault(discretization_key(q.problem), []).append(q)
        for reqs in by.values():
            cls = self._class(reqs[0].problem)
            ens = Ensemble.from_problems([q.problem for q in reqs], self.scheme, self.dt, mode=cls.mode)
            ens.set_states(np.stack([q.X0 for q in reqs]))
            self.startups.append((ens, reqs, torch.cuda.Stream()))

    def _join_ready(self) -> None:
        """Startup cohorts whose history is complete move their unfinished runs into slots."""
        keep = []
        for ens, reqs, st in self.startups:
            im = ens.imex
            if im.history_length < im.order - 1:
                keep.append((ens, reqs, st))
                continue
            cls = self._class(reqs[0].problem)
            fac = _full_order_factor(im)
            for b, q in enumerate(reqs):
                if q.result is None:
                    cls.join(ens, b, q, fac)
                    self.joins += 1
            _torch().cuda.current_stream().wait_stream(cls.stream)
        self.startups = keep

    def tick(self, now: float) -> float:
        """Admit, join, step every class and every startup cohort once, retire; returns the
        tick's wall time (s)."""
        torch = _torch()
        t0 = time.perf_counter()
        self._admit(now)
        self._join_ready()
        active = [c for c in self.classes.values() if c.members]
        for c in active:
            c.step()
        for ens, reqs, st in self.startups:
            with torch.cuda.stream(st):
                ens.run(1)
        torch.cuda.synchronize()
        lat = time.perf_counter() - t0
        for c in active:
            done = []
            for s, q in c.members.items():
                q.step_latencies.append(lat)
                if len(q.step_latencies) >= q.steps:
                    done.append(s)
            if done:
                X = c.ens.states()
                for s in done:
                    self._retire(c.members.pop(s), X[s], now + lat)
        for ens, reqs, st in self.startups:
            live = [(b, q) for b, q in enumerate(reqs) if q.result is None]
            for b, q in live:
                q.step_latencies.append(lat)
            if any(len(q.step_latencies) >= q.steps for b, q in live):
                X = ens.states()
                for b, q in live:
                    if len(q.step_latencies) >= q.steps:
                        self._retire(q, X[b], now + lat)
        self.startups = [s for s in self.startups if any(q.result is None for q in s[1])]
        self.ticks.append(lat)
        return lat

    def _retire(self, q: Request, X: np.ndarray, t: float) -> None:
        q.result = np.array(X)
        q.finished = t
        self.finished.append(q)
        self.resident -= footprint_bytes(q.problem, self._order)

    def _busy(self) -> bool:
        return bool(self.startups) or any(c.members for c in self.classes.values())

    def serve(self, requests: list[Request], clock: str = "wall") -> dict:
        """Submit ``requests`` (their ``arrival`` times) and serve until all finished.
        clock='wall': a request becomes admissible when the wall time since the start passes its
        arrival in seconds (idle gaps are slept); clock='ticks': arrivals count in ticks."""
        for q in sorted(requests, key=lambda q: q.arrival):
            self.submit(q)
        start = time.perf_counter()
        ntick = 0
        while self.waiting or self._busy():
            now = (time.perf_counter() - start) if clock == "wall" else float(ntick)
            if not self._busy() and self.waiting and min(q.arrival for q in self.waiting) > now:
                nxt = min(q.arrival for q in self.waiting)
                if clock == "wall":
                    time.sleep(nxt - now)
                else:
                    ntick = int(np.ceil(nxt))
                continue
            self.tick(now)
            ntick += 1
        wall = time.perf_counter() - start
        return self.stats(wall)

    def stats(self, wall: float) -> dict:
        lat = np.array([x for q in self.finished for x in q.step_latencies]) * 1e3
        steps = int(lat.size)
        return {"requests": len(self.finished), "steps": steps, "wall_s": wall,
                "steps_per_s": steps / wall if wall > 0 else float("nan"),
                "p50_step_ms": float(np.percentile(lat, 50)) if steps else float("nan"),
                "p99_step_ms": float(np.percentile(lat, 99)) if steps else float("nan"),
                "ticks": len(self.ticks), "joins": self.joins,
                "slots": {f"{k[0]}x{k[1]}": c.capacity for k, c in self.classes.items()},
                "modes": {f"{k[0]}x{k[1]}": v for k, v in self._modes.items()}}
