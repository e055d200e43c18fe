"""Serving of IVP runs on one MI355X: continuous batching of 2D Rayleigh-Benard requests.

Requests (a problem, an initial state, a number of steps) arrive at any time.  The server keeps
every admitted run's state, SBDF history and factors resident in HBM and advances all of them in
ticks; one tick is one step of every active run:

* admission, at the start of a tick: waiting requests are admitted in arrival order while the
  estimated HBM footprint of all admitted runs stays within the budget; the requests admitted in
  one tick start together, one cohort per size class (an ``Ensemble``: one batched explicit
  term, stage and solve for all its members), so that new work joins at a step boundary without
  waiting for the runs already on the device;
* stepping: every cohort makes one step on its own HIP stream (full-order steps replayed from the
  cohort's captured graphs), so that the cohorts' kernels overlap on the device; the tick ends
  when the device has finished all of them (that wall time is the step latency of every run that
  advanced in the tick);
* retirement: a run that reached its step count leaves its cohort with its final state copied to
  the host; the cohort is then compacted to its remaining members, whose states and SBDF history
  levels are carried over (``Imex.adopt_history``);
* merging: cohorts of one size class that have all reached full order become one cohort (their
  states and history levels copied, ring phases mapped), so that runs admitted at different
  ticks share one batched step from then on.
Compaction and merging only copy values: every run's arithmetic stays the one of its single run
(tested: each result is the same alone and in the mix).

The kernel mode of each size class is picked once by ``autotune_mode``.  There is no counterpart
in the reference (it has no server); the per-run step is the reference's IVPSolver step
(python/conatus/timestep/integrator.py:648-682).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .ensembles import Ensemble, autotune_mode, discretization_key
from .errors import OptionError
from .problems import RayleighBenard2D

__all__ = ["Request", "Server", "footprint_bytes"]


def _torch():
    import torch
    return torch


@dataclass
class Request:
    problem: RayleighBenard2D
    steps: int
    X0: np.ndarray
    arrival: float = 0.0                  # seconds after the start of the serving loop
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


class _Cohort:
    def __init__(self, requests: list, scheme: str, dt: float, mode: str | None):
        torch = _torch()
        self.requests = requests
        self.ens = Ensemble.from_problems([q.problem for q in requests], scheme, dt, mode=mode)
        self.ens.set_states(np.stack([q.X0 for q in requests]))
        self.stream = torch.cuda.Stream()
        self.done_steps = 0

    def step(self) -> None:
        torch = _torch()
        with torch.cuda.stream(self.stream):
            self.ens.run_graph(1)
        self.done_steps += 1

    def full_order(self) -> bool:
        im = self.ens.imex
        return not im.order or im.history_length == im.order

    def compact(self, keep: list) -> None:
        """Keep members ``keep`` (indices), carrying their states and history over."""
        merged = _Cohort.merge([(self, keep)])
        self.ens, self.requests = merged.ens, merged.requests

    @staticmethod
    def merge(parts: list) -> "_Cohort":
        """One cohort of the members ``keep`` of every (cohort, keep) in ``parts`` (same size
        class, same history length): states and SBDF history levels are copied, so that every
        member's next step is the one its own cohort would have made."""
        torch = _torch()
        c0 = parts[0][0]
        old0 = c0.ens
        out = _Cohort.__new__(_Cohort)
        out.stream = c0.stream
        out.requests = [c.requests[b] for c, keep in parts for b in keep]
        probs = [c.ens.problems[b] for c, keep in parts for b in keep]
        K, n = old0.K, old0.n
        with torch.cuda.stream(out.stream):
            for c, _ in parts:
                out.stream.wait_stream(c.stream)
            new = Ensemble.from_problems(probs, old0.scheme, old0.dt, mode=old0.mode)
            sources = []
            for c, keep in parts:
                rows = torch.cat([torch.arange(b * K, (b + 1) * K, device="cuda") for b in keep])
                sources.append((c.ens.imex, rows))
            new._flat().copy_(torch.cat([c.ens._flat().index_select(0, rows) for (c, _), (_, rows) in zip(parts, sources)]))
            new.t = old0.t
            new.imex.adopt_history(sources, new._flat(), new.t)
        out.ens = new
        out.done_steps = c0.done_steps
        return out


class Server:
    """Continuous-batching server of 2D Rayleigh-Benard IVP requests (module doc)."""

    def __init__(self, scheme: str = "SBDF2", dt: float = 1e-4, hbm_budget: float = 200e9,
                 autotune: bool = True, mode: str | None = None, merge: bool = True):
        if autotune and mode is not None:
            raise OptionError("Server: autotune and a fixed mode exclude each other")
        self.scheme, self.dt = scheme, float(dt)
        self.budget = float(hbm_budget)
        self.autotune, self.mode = autotune, mode
        self._modes: dict = {}
        self.waiting: list[Request] = []
        self.cohorts: list[_Cohort] = []
        self.finished: list[Request] = []
        self.resident = 0
        self.ticks: list[float] = []
        self._order = {"SBDF1": 1, "SBDF2": 2, "SBDF3": 3, "SBDF4": 4}.get(scheme, 0)
        self._next_id = 0
        self.merge = merge
        self.merges = 0

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

    # ------------------------------------------------------------------ scheduling
    def _admit(self, now: float) -> None:
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
        classes: dict = {}
        for q in take:
            classes.setdefault(discretization_key(q.problem), []).append(q)
        for key, reqs in classes.items():
            self.cohorts.append(_Cohort(reqs, self.scheme, self.dt, self._mode_for(reqs[0].problem)))

    def tick(self, now: float) -> float:
        """Admit, step every cohort once, retire; returns the tick's wall time (s)."""
        torch = _torch()
        t0 = time.perf_counter()
        self._admit(now)
        for c in self.cohorts:
            c.step()
        torch.cuda.synchronize()
        lat = time.perf_counter() - t0
        alive: list[_Cohort] = []
        for c in self.cohorts:
            keep = []
            for q in c.requests:
                q.step_latencies.append(lat)
            states = None
            for b, q in enumerate(c.requests):
                if len(q.step_latencies) >= q.steps:
                    if states is None:
                        states = c.ens.states()
                    q.result = states[b]
                    q.finished = now + lat
                    self.finished.append(q)
                    self.resident -= footprint_bytes(q.problem, self._order)
                else:
                    keep.append(b)
            if keep and len(keep) < len(c.requests):
                c.compact(keep)
            if keep:
                alive.append(c)
        self.cohorts = self._merge(alive)
        self.ticks.append(lat)
        return lat

    def _merge(self, cohorts: list) -> list:
        """Cohorts of one size class that all run at full order become one (continuous batching:
        runs admitted at different ticks share one batched step from then on)."""
        if not self.merge:
            return cohorts
        by: dict = {}
        out = []
        for c in cohorts:
            if c.full_order():
                by.setdefault(discretization_key(c.ens.problems[0]), []).append(c)
            else:
                out.append(c)
        for key, cs in by.items():
            if len(cs) == 1:
                out.append(cs[0])
            else:
                out.append(_Cohort.merge([(c, list(range(len(c.requests)))) for c in cs]))
                self.merges += 1
        return out

    def serve(self, requests: list[Request], clock: str = "wall") -> dict:
        """Submit ``requests`` (their ``arrival`` times in seconds) and serve until all finished.
        clock='wall': a request becomes admissible when the wall time since the start passes its
        arrival (idle gaps are slept); clock='ticks': arrivals count in ticks (tests)."""
        for q in sorted(requests, key=lambda q: q.arrival):
            self.submit(q)
        start = time.perf_counter()
        ntick = 0
        while self.waiting or self.cohorts:
            now = (time.perf_counter() - start) if clock == "wall" else float(ntick)
            if not self.cohorts and self.waiting and min(q.arrival for q in self.waiting) > now:
                if clock == "wall":
                    time.sleep(min(q.arrival for q in self.waiting) - now)
                    continue
                ntick = int(np.ceil(min(q.arrival for q in self.waiting)))
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
                "ticks": len(self.ticks), "merges": self.merges, "modes": {f"{k[0]}x{k[1]}": v for k, v in self._modes.items()}}
