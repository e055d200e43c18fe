"""Serving of IVP runs on one MI355X: continuous batching of 2D Rayleigh-Benard and 3D channel
requests.

Requests (a problem, an initial state, a number of steps) arrive at any time.  The server keeps
every admitted run's state, SBDF history and factored mode systems resident in HBM and advances
them page by page, each page (a batch of runs of one size class) at its own pace.

* Pages: every size class (grid) holds pages, resident ``Ensemble``s of ``page_size`` slots
  (``channel_page_size`` for channels, whose steps are larger) built once (explicit-term tables,
  CSR, factors, captured graphs).  A stepping page runs at full order
  and is replayed from its graphs (one batched explicit term, stage and solve for its slots); a
  startup page makes the lower-order startup steps of the runs placed in it.  Only pages that
  hold runs are stepped, each on its own stream, so that the pages' kernels overlap.  The kernel
  mode of a class is picked once by ``autotune_mode``.
* Admission, at the start of a tick: waiting requests are admitted in arrival order while the
  estimated HBM footprint of the admitted runs stays within the budget, into free slots of a
  startup page of their class (state copied in).
* Factors: a slot holds the factored operators of its run (one per SBDF order it steps at),
  copied from the run's single-run handles (``cn_bb_copy_systems``), which are built once per
  (grid, parameters, order) and cached; a slot that already holds the same parameters (Ra, Pr of
  RB 2D, Re of a channel) copies nothing.
* Joining: a run whose startup is complete moves into a free slot of a stepping page of its class
  (the fullest first) by device copies only: its state and its SBDF history levels (into the
  ring slots the page's next step reads).  From then on it shares the page's batched step.
* Retirement: a run that reached its step count leaves with its state copied to the host; its
  slot is free for the next run (a free slot keeps being stepped, its values are never read).
* Clocks: every page steps on its own stream at its own pace.  The serving loop issues one step
  of every idle page that holds runs (recording an event behind it), then polls the events: a
  page whose step has completed counts that step for the runs it held when the step was issued,
  retires the runs that are done, and is issued again.  Nothing waits for the whole device, so a
  small run's steps are not held up by a large run's.  A run's step latency is the wall time from
  the issue of its page's step to the moment the loop sees it complete.  With clock='ticks' the
  loop instead waits for every issued step before the next round (one round = one tick), which
  makes admission times reproducible in tests; the arithmetic is the same either way.
* Hardware queues: HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
  default); with more stepping pages than queues, a small page's step can wait in a shared queue
  behind a large page's.  A serving process sets GPU_MAX_HW_QUEUES (16 in the benchmark: per-class
  p50 0.3-1.4 ms instead of 2-3 ms for the RB classes) before its first HIP call.

Moving a run only copies values: every run's arithmetic stays the one of its single run (tested:
each result is the same alone and in the mix).  There is no counterpart in the reference (it has
no server); the per-run step is the reference's IVPSolver step
(python/conatus/timestep/integrator.py:648-682).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .ensembles import Ensemble, autotune_mode, discretization_key
from .errors import OptionError
from .problems import RayleighBenard2D
from .runtime.imex import factor_operator, sbdf_coefficients

__all__ = ["Request", "Server", "footprint_bytes"]


def _torch():
    import torch
    return torch


@dataclass
class Request:
    problem: object                       # RayleighBenard2D or Channel3D
    steps: int
    X0: np.ndarray
    arrival: float = 0.0                  # seconds (or ticks) after the start of the serving loop
    rid: int = -1
    # filled by the server
    admitted: float | None = None
    finished: float | None = None
    result: np.ndarray | None = field(default=None, repr=False)
    step_latencies: list = field(default_factory=list, repr=False)


def footprint_bytes(pb, order: int = 2) -> int:
    """HBM bytes of one resident run under SBDF-``order``: the device-memory plan of its solver
    (conatus_amd.memory.plan: state, history, operators, the factor handles of every startup
    order, explicit-term tables and workspace, the measured runtime share)."""
    from .memory import plan
    return plan(pb, f"SBDF{int(order)}").total


def _params(pb) -> tuple:
    return pb.params()


def _grid(key: tuple) -> str:
    """'NxxNz' of a RB 2D discretization key, 'channel NxxNyxNz' of a channel's."""
    if key and key[0] == "channel3d":
        return f"channel {key[1]}x{key[2]}x{key[3]}"
    return f"{key[0]}x{key[1]}"


def _alpha(order: int, dt: float) -> float:
    """The implicit coefficient a0/dt of an SBDF step of the given order."""
    return sbdf_coefficients(order)[0][0] / dt


class _Page:
    """A resident ensemble of ``P`` slots of one size class (the template's problem in every slot).
    A stepping page runs at full order from the start (runs join with their own history) and is
    replayed from its captured graphs; a startup page makes the lower-order startup steps of the
    runs placed in it, eagerly.  ``facs`` maps an SBDF order to the page's factored operator of
    that order; ``slot_params[q][s]`` is the (Ra, Pr) whose factors slot s holds for order q."""

    def __init__(self, template, P: int, scheme: str, dt: float, mode: str | None,
                 startup: bool):
        torch = _torch()
        self.P, self.startup, self.dt = P, startup, dt
        self.stream = torch.cuda.Stream()
        self.event = torch.cuda.Event()
        self.members: dict = {}
        self.inflight: dict | None = None     # the members of the step in flight (None: idle)
        self.t_issue = 0.0
        self.ens = Ensemble.from_problems([template] * P, scheme, dt, mode=mode)
        im = self.ens.imex
        with torch.cuda.stream(self.stream):
            orders = range(1, im.order) if startup else [im.order]
            self.facs = {q: im._factor(_alpha(q, dt), 1.0) for q in orders}
            if not startup:
                im._levels = list(range(im.order))
                for _ in range(im.order):            # capture every ring phase's graph now
                    im.replay_step(self.ens._flat())
        self.slot_params = {q: [_params(template)] * P for q in self.facs}

    def free_slots(self) -> list:
        return [s for s in range(self.P) if s not in self.members]

    def set_factors(self, s: int, req: Request, runfac) -> None:
        """Slot s gets the factors of req's operators (copied from its single-run handles)."""
        K = self.ens.K
        for q, fac in self.facs.items():
            if self.slot_params[q][s] != _params(req.problem):
                fac.copy_systems_from(runfac(req.problem, q), 0, s * K, K, self.stream.cuda_stream)
                self.slot_params[q][s] = _params(req.problem)

    def step(self) -> None:
        """Issue one step of every slot on the page's stream (no host synchronisation)."""
        torch = _torch()
        with torch.cuda.stream(self.stream):
            if self.startup:
                self.ens.imex.startup_step(self.ens._flat(), self.ens.t)
            else:
                self.ens.imex.replay_step(self.ens._flat())
            self.ens.t += self.ens.dt

    def issue(self) -> None:
        self.inflight = dict(self.members)
        self.t_issue = time.perf_counter()
        self.step()
        self.event.record(self.stream)

    def slot_state(self, s: int) -> np.ndarray:
        """Host copy of slot s (waits for the page's stream only)."""
        torch = _torch()
        K = self.ens.K
        with torch.cuda.stream(self.stream):
            return self.ens._flat()[s * K:(s + 1) * K].cpu().numpy()


class Server:
    """Continuous-batching server of 2D Rayleigh-Benard and 3D channel IVP requests (module doc)."""

    def __init__(self, scheme: str = "SBDF2", dt: float = 1e-4, hbm_budget: float = 200e9,
                 autotune: bool = True, mode: str | None = None, page_size: int = 8,
                 channel_page_size: int = 2):
        if autotune and mode is not None:
            raise OptionError("Server: autotune and a fixed mode exclude each other")
        if scheme not in ("SBDF1", "SBDF2", "SBDF3", "SBDF4"):
            raise OptionError(f"Server: scheme {scheme!r} not served", fix="one of SBDF1..SBDF4")
        for name, v in (("page_size", page_size), ("channel_page_size", channel_page_size)):
            if int(v) != v or v < 1:
                raise OptionError(f"Server: {name} must be a whole number >= 1")
        self.channel_page_size = int(channel_page_size)
        self.scheme, self.dt = scheme, float(dt)
        self.budget = float(hbm_budget)
        self.autotune, self.mode = autotune, mode
        self.page_size = int(page_size)
        self._order = int(scheme[-1])
        self._modes: dict = {}
        self.pages: dict = {}             # (grid, startup) -> [_Page]
        self._runfacs: dict = {}
        self.pages_built = 0
        self.waiting: list[Request] = []
        self.finished: list[Request] = []
        self.resident = 0
        self.page_steps: list[float] = []   # wall time of every completed page step
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

    def _mode_for(self, pb) -> str | None:
        if not self.autotune:
            return self.mode
        key = discretization_key(pb)
        if key not in self._modes:
            # timed on a full page of the class: the mode that solves one run fastest (few systems:
            # the partitioned sweeps) is not the one for a page of them
            P = self.page_size if isinstance(pb, RayleighBenard2D) else self.channel_page_size
            self._modes[key] = autotune_mode([pb] * P, self.scheme, self.dt)[0]
        return self._modes[key]

    def _runfac(self, pb, q: int):
        """Factored operator of one run's K mode systems for SBDF order q, cached per (grid,
        parameters, q): the source of the factors copied into slots."""
        key = (discretization_key(pb), _params(pb), q)
        f = self._runfacs.get(key)
        if f is None:
            Ms = [M.astype(np.complex128) for M, L in pb.mats]
            Ls = [L.astype(np.complex128) for M, L in pb.mats]
            f = factor_operator(Ms, Ls, _alpha(q, self.dt), 1.0, pb.border, np.complex128, True,
                                self._mode_for(pb))
            self._runfacs[key] = f
        return f

    def _pages(self, pb, startup: bool) -> list:
        return self.pages.setdefault((discretization_key(pb), startup), [])

    def _new_page(self, pb, startup: bool) -> _Page:
        P = self.page_size if isinstance(pb, RayleighBenard2D) else self.channel_page_size
        pg = _Page(pb, P, self.scheme, self.dt, self._mode_for(pb), startup)
        self._pages(pb, startup).append(pg)
        self.pages_built += 1
        return pg

    def prepare(self, problems, pages: int = 2, startup_pages: int = 1) -> None:
        """Tune the size classes of ``problems`` and build their pages (``pages`` stepping pages
        and ``startup_pages`` startup pages each) and the single-run factors of their parameters,
        ahead of the requests (a page built while serving costs its tick the setup: CSR
        operators, factorization and graph capture of all its slots)."""
        for pb in problems:
            while len(self._pages(pb, False)) < pages:
                self._new_page(pb, False)
            while self._order > 1 and len(self._pages(pb, True)) < startup_pages:
                self._new_page(pb, True)
            for q in range(1, self._order + 1):
                self._runfac(pb, q)

    # ------------------------------------------------------------------ scheduling
    def _admit(self, now: float) -> None:
        """Admit waiting requests (arrival order, HBM footprint) into free startup slots."""
        torch = _torch()
        rest: list[Request] = []
        blocked = False
        for q in self.waiting:
            if q.arrival > now:
                rest.append(q)
                continue
            fp = footprint_bytes(q.problem, self._order)
            if blocked or self.resident + fp > self.budget:
                blocked = True          # arrival order: nothing overtakes a request that waits
                rest.append(q)
                continue
            self.resident += fp
            q.admitted = now
            if self._order == 1:
                self._join(q, None, 0)
                continue
            # a startup page that has not started stepping (no member has stepped yet)
            pg = next((p for p in self._pages(q.problem, True)
                       if p.free_slots() and not any(m.step_latencies for m in p.members.values())), None)
            if pg is None:
                pg = next((p for p in self._pages(q.problem, True) if not p.members), None) \
                    or self._new_page(q.problem, True)
            s = pg.free_slots()[0]
            if not pg.members:
                pg.ens.imex.reset()
                pg.ens.t = 0.0
            self._upload(pg, s, q)
            pg.members[s] = q
        self.waiting = rest

    def _upload(self, pg: _Page, s: int, q: Request) -> None:
        """Initial state and factors of run q into slot s of page pg, on the page's stream: the
        state travels from a pinned host copy (kept with the request until the run retires), so
        that the host does not wait for the copy."""
        torch = _torch()
        K = pg.ens.K
        q._pinned = torch.from_numpy(np.ascontiguousarray(q.X0, dtype=np.complex128)).pin_memory()
        with torch.cuda.stream(pg.stream):
            pg.ens._flat()[s * K:(s + 1) * K].copy_(q._pinned, non_blocking=True)
            pg.set_factors(s, q, self._runfac)

    def _join(self, q: Request, src, sslot: int) -> None:
        """Move run q (state and history in slot sslot of startup page src, or its initial state
        when the scheme has no startup) into a free slot of a stepping page of its class."""
        torch = _torch()
        pages = self._pages(q.problem, False)
        pg = max((p for p in pages if p.free_slots()), key=lambda p: len(p.members), default=None) \
            or self._new_page(q.problem, False)
        s = pg.free_slots()[0]
        im, K, n = pg.ens.imex, pg.ens.K, pg.ens.n
        with torch.cuda.stream(pg.stream):
            if src is not None:
                pg.stream.wait_stream(src.stream)
                pg.ens._flat()[s * K:(s + 1) * K].copy_(src.ens._flat()[sslot * K:(sslot + 1) * K])
                cim = src.ens.imex
                # history level i of the run goes to the ring slot that is level i of the page
                # now: the page's next step reads it as level i + 1
                for i, cs in enumerate(cim._levels):
                    for j in range(2):
                        im._ring[im._levels[i]][j].view(im.batch, n)[s * K:(s + 1) * K].copy_(
                            cim._ring[cs][j].view(cim.batch, n)[sslot * K:(sslot + 1) * K])
                pg.set_factors(s, q, self._runfac)
        if src is None:
            self._upload(pg, s, q)
        if src is not None:
            # the startup page may be reset and refilled in this same tick (admission writes its
            # slots on its own stream): that work waits until the copies above have read them
            src.stream.wait_stream(pg.stream)
        pg.members[s] = q
        self.joins += 1

    def _join_ready(self) -> None:
        """Idle startup pages whose history is complete move their runs into stepping pages."""
        for (key, startup), pages in list(self.pages.items()):
            if not startup:
                continue
            for pg in pages:
                im = pg.ens.imex
                if pg.members and pg.inflight is None and im.history_length >= im.order - 1:
                    for s, q in sorted(pg.members.items()):
                        self._join(q, pg, s)
                    pg.members = {}

    def _all_pages(self) -> list:
        return [pg for pages in self.pages.values() for pg in pages]

    def _issue_idle(self) -> int:
        """Issue one step of every idle page that holds runs; returns how many were issued."""
        k = 0
        for pg in self._all_pages():
            if pg.members and pg.inflight is None:
                pg.issue()
                k += 1
        return k

    def _complete(self, pg: _Page, now: float) -> None:
        """Count the completed step of page pg for the runs it held at issue; retire finished runs."""
        lat = time.perf_counter() - pg.t_issue
        stepped, pg.inflight = pg.inflight, None
        self.page_steps.append(lat)
        for s, q in stepped.items():
            q.step_latencies.append(lat)
            if len(q.step_latencies) >= q.steps and pg.members.get(s) is q:
                self._retire(pg.members.pop(s), pg.slot_state(s), now)

    def _poll(self, now: float) -> int:
        """Complete every page whose step has finished on the device; returns how many did."""
        k = 0
        for pg in self._all_pages():
            if pg.inflight is not None and pg.event.query():
                self._complete(pg, now)
                k += 1
        return k

    def tick(self, now: float) -> float:
        """One synchronous round: join, admit, step every page that holds runs and wait for all of
        them, retire; returns the round's wall time (s)."""
        t0 = time.perf_counter()
        self._join_ready()
        self._admit(now)
        self._issue_idle()
        for pg in self._all_pages():
            if pg.inflight is not None:
                pg.event.synchronize()
                self._complete(pg, now)
        return time.perf_counter() - t0

    def _retire(self, q: Request, X: np.ndarray, t: float) -> None:
        q.result = np.array(X)
        q.finished = t
        q._pinned = None
        self.finished.append(q)
        self.resident -= footprint_bytes(q.problem, self._order)

    def _busy(self) -> bool:
        return any(pg.members or pg.inflight is not None for pg in self._all_pages())

    def serve(self, requests: list[Request], clock: str = "wall") -> dict:
        """Submit ``requests`` (their ``arrival`` times) and serve until all finished.
        clock='wall': a request becomes admissible when the wall time since the start passes its
        arrival in seconds (idle gaps are slept); clock='ticks': arrivals count in ticks."""
        if clock not in ("wall", "ticks"):
            raise OptionError(f"Server.serve: clock={clock!r}", fix="clock='wall' or clock='ticks'")
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
            if clock == "ticks":
                self.tick(now)
                ntick += 1
                continue
            # wall clock: pages step independently; the loop reacts to whichever finishes first
            self._poll(now)
            self._join_ready()
            self._admit(now)
            if self._issue_idle() == 0 and not self._poll(time.perf_counter() - start):
                time.sleep(2e-6)
        wall = time.perf_counter() - start
        return self.stats(wall)

    def stats(self, wall: float) -> dict:
        lat = np.array([x for q in self.finished for x in q.step_latencies]) * 1e3
        steps = int(lat.size)
        by_class: dict = {}
        for q in self.finished:
            by_class.setdefault(_grid(discretization_key(q.problem)), []).extend(q.step_latencies)
        per_class = {g: {"requests": sum(1 for q in self.finished if _grid(discretization_key(q.problem)) == g),
                         "steps": len(v), "p50_step_ms": float(np.percentile(v, 50)) * 1e3,
                         "p99_step_ms": float(np.percentile(v, 99)) * 1e3}
                     for g, v in sorted(by_class.items()) if v}
        return {"requests": len(self.finished), "steps": steps, "wall_s": wall,
                "steps_per_s": steps / wall if wall > 0 else float("nan"),
                "p50_step_ms": float(np.percentile(lat, 50)) if steps else float("nan"),
                "p99_step_ms": float(np.percentile(lat, 99)) if steps else float("nan"),
                "per_class": per_class,
                "page_steps": len(self.page_steps), "joins": self.joins, "pages_built": self.pages_built,
                "pages": {f"{_grid(k[0])}{' startup' if k[1] else ''}": len(v) for k, v in self.pages.items()},
                "modes": {_grid(k): v for k, v in self._modes.items()}}
