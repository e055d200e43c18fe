"""Checkpoints of an IVP run: state, multistep history and clock, for a bit-for-bit restart.

The file is a ``conatus.checkpoint/1`` NumPy payload as the reference writes it
(python/conatus/timestep/integrator.py:974-1008, python/conatus/io/checkpoint.py:116-142): a
``meta`` JSON record and the arrays ``X`` (the stacked state, one vector), ``MX<i>`` and ``F<i>``
(``nhist`` multistep levels, newest first: M X and F(X) of earlier states) and, for the
Crank-Nicolson family, ``LX<i>`` (``nhist_LX`` levels of L X).  The record carries the
reference's keys (scheme, dt, startup, startup_substeps, n, iteration, t, t_origin, it_origin,
reduced, ordering, problem, nhist, nhist_LX, bc_time, time_parameter, history_origin,
t_initial, identity) plus ``format``.  The stacked state is this package's mode ordering
(``ordering`` names it: unknown j of Fourier mode k at index k n + j), so a file restarts the
solver that wrote it or one of the same discretization; the reference's own stacking differs.

Restart refuses other numerics (scheme, dt, startup, n, ordering, pivoting) and another problem
(grid, walls, dealiasing, parameters), and checks that every array the record declares is
present with the state's shape before anything of the solver changes (the reference's
check_payload).  A record without ``format`` (the reference's ``.npz`` meta has none) is read
the same way.  The ``.npz`` is written to a temporary name and renamed, so that a job killed
while writing never leaves a truncated checkpoint; HDF5 needs h5py, which is absent: ``.h5``
paths are refused.  Servers use the in-memory payload (``payload`` / ``restore``) to evict and
resume runs.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from .errors import OptionError

__all__ = ["FORMAT", "ORDERING", "payload", "restore", "save_checkpoint", "load_checkpoint"]

FORMAT = "conatus.checkpoint/1"
ORDERING = "conatus_amd.modes/1"          # stacked state: unknown j of mode k at k * n + j


def _levels(im) -> list[tuple]:
    """The history levels of the integrator, newest first: (M X, F) for SBDF, (M X, F, L X) for
    the CN family; none for Runge-Kutta schemes."""
    if not (im.order or im.cn is not None):
        return []
    return [im._ring[slot] for slot in im._levels]


def payload(solver) -> tuple[dict, dict]:
    """(meta, host arrays) of an IVPSolver: stacked state, multistep history (newest first), clock."""
    from . import __version__
    im = solver.imex
    levels = _levels(im)
    cn = im.cn is not None
    pb = solver.problem
    meta = {"format": FORMAT, "scheme": solver.scheme, "dt": float(solver.dt), "startup": "lower_order",
            "startup_substeps": 1, "n": int(pb.K * pb.n), "iteration": int(im.iteration),
            "t": float(solver.t), "t_origin": float(solver.t_origin), "it_origin": int(solver.it_origin),
            "reduced": False, "ordering": ORDERING, "problem": pb.describe(), "nhist": len(levels),
            "nhist_LX": len(levels) if cn else 0, "bc_time": "step", "time_parameter": None,
            "history_origin": int(im.iteration) - len(levels), "t_initial": float(solver.t_origin),
            "pivot": bool(solver.pivot), "conatus_version": f"conatus_amd {__version__}",
            "identity": pb.identity()}
    arrays = {"X": solver.state().reshape(-1)}
    for i, lv in enumerate(levels):
        arrays[f"MX{i}"] = lv[0].cpu().numpy().reshape(-1)
        arrays[f"F{i}"] = lv[1].cpu().numpy().reshape(-1)
        if cn:
            arrays[f"LX{i}"] = lv[2].cpu().numpy().reshape(-1)
    return meta, arrays


def _missing(meta: dict, arrays: dict, nstate: int) -> str | None:
    """What is absent or misshapen among the arrays the record declares (None: complete)."""
    try:
        nh, nl = int(meta["nhist"]), int(meta.get("nhist_LX", 0))
    except (KeyError, TypeError, ValueError):
        return "the record has no history count (nhist)"
    want = ["X"] + [f"{k}{i}" for i in range(nh) for k in ("MX", "F")] + [f"LX{i}" for i in range(nl)]
    absent = [k for k in want if k not in arrays]
    if absent:
        return "missing " + ", ".join(absent)
    wrong = [f"{k} has shape {np.shape(arrays[k])}" for k in want if np.shape(arrays[k]) != (nstate,)]
    if wrong:
        return "; ".join(wrong) + f" (the state has {nstate} entries)"
    return None


def restore(solver, meta: dict, arrays: dict, where: str = "checkpoint") -> None:
    """Continue ``solver`` from a payload.  Every check runs before the solver changes."""
    torch = __import__("torch")
    fmt = meta.get("format", FORMAT)
    if fmt != FORMAT:
        raise OptionError(f"{where}: not a Conatus checkpoint (format {fmt!r}, expected {FORMAT!r})")
    pb, im = solver.problem, solver.imex
    mine = {"scheme": solver.scheme, "dt": float(solver.dt), "startup": "lower_order", "startup_substeps": 1,
            "n": int(pb.K * pb.n), "reduced": False, "ordering": ORDERING}
    for key, have in mine.items():
        if key not in meta:
            raise OptionError(f"{where}: the record has no {key!r}", fix="restart from a complete checkpoint")
        if meta[key] != have:
            raise OptionError(f"{where}: checkpoint mismatch on {key}: {meta[key]!r} vs {have!r} "
                              "(a restart never changes the numerics silently)",
                              fix=f"build the solver with {key}={meta[key]!r} to restart from this checkpoint")
    if "pivot" in meta and bool(meta["pivot"]) != bool(solver.pivot):
        raise OptionError(f"{where}: checkpoint mismatch on pivot: {meta['pivot']!r} vs {solver.pivot!r}")
    ident = pb.identity()
    if meta.get("identity") != ident:
        theirs = meta.get("identity") or {}
        diff = sorted(k for k in set(ident) | set(theirs) if theirs.get(k) != ident.get(k))
        raise OptionError(f"{where}: the checkpoint is a state of another problem (differs in {', '.join(diff)})")
    N = pb.K * pb.n
    problem = _missing(meta, arrays, N)
    if problem:
        raise OptionError(f"{where} is an incomplete checkpoint: {problem}",
                          why="a restart needs the state and the whole history its record declares; "
                              "the solver was left unchanged",
                          fix="restart from an earlier complete checkpoint")
    nh, nl = int(meta["nhist"]), int(meta.get("nhist_LX", 0))
    cn = im.cn is not None
    cap = (im.order if im.order else len(im._ring)) if (im.order or cn) else 0
    if nh > cap or (cn and nl != nh) or (not cn and nl):
        raise OptionError(f"{where}: {nh} M X / F and {nl} L X history levels do not fit {solver.scheme}")
    # checks done: the solver changes from here on
    solver.X.copy_(torch.from_numpy(np.ascontiguousarray(np.asarray(arrays["X"]).reshape(pb.K, pb.n),
                                                          dtype=np.complex128)))
    im.reset()
    for i in range(nh):
        parts = ("MX", "F", "LX") if cn else ("MX", "F")
        for j, key in enumerate(parts):
            im._ring[i][j].copy_(torch.from_numpy(np.ascontiguousarray(arrays[f"{key}{i}"], dtype=np.complex128)))
    im._levels = list(range(nh))
    im.iteration = int(meta["iteration"])
    solver.t = float(meta["t"])
    solver.t_origin = float(meta.get("t_origin", meta["t"]))
    solver.it_origin = int(meta.get("it_origin", meta["iteration"] if "t_origin" not in meta else 0))
    im._remember(solver.X, solver.t)


def save_checkpoint(solver, path: str | Path) -> Path:
    """Write the payload of ``solver`` to ``path`` (``.npz``; the suffix is added if missing).
    A non-finite state is never written (the solver's finite-state check runs first)."""
    path = Path(path)
    if path.suffix in (".h5", ".hdf5"):
        raise OptionError("save_checkpoint: HDF5 checkpoints need h5py, which is not available",
                          fix="save_checkpoint('state.npz')")
    if getattr(solver, "_finite_every", 0):
        solver.check_finite(f"save_checkpoint({path.name})")
    if not str(path).endswith(".npz"):
        path = Path(str(path) + ".npz")
    meta, arrays = payload(solver)
    part = path.with_name(path.name[:-4] + ".part.npz")
    np.savez(part, meta=json.dumps(meta), **arrays)
    os.replace(part, path)
    return path


def load_checkpoint(solver, path: str | Path) -> None:
    """Continue ``solver`` from a ``conatus.checkpoint/1`` ``.npz`` (pickled objects refused)."""
    path = Path(path)
    if not path.exists() and path.with_suffix(".npz").exists():
        path = path.with_suffix(".npz")
    if not path.exists():
        raise OptionError(f"load_checkpoint: {path} does not exist")
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    if "meta" not in arrays:
        raise OptionError(f"load_checkpoint: {path} has no meta record (not a Conatus checkpoint)")
    meta = json.loads(str(arrays.pop("meta")))
    restore(solver, meta, arrays, where=str(path))
