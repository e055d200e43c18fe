"""Checkpoints of an IVP run: state, SBDF history and clock, for a bit-for-bit restart.

The reference writes ``conatus.checkpoint/1`` payloads (python/conatus/timestep/integrator.py:974-1076,
python/conatus/io/checkpoint.py:1-118): a ``meta`` JSON record (scheme, dt, iteration, t, history
length, what the state is a state of) and the arrays ``X`` (state), ``MX<i>`` and ``F<i>``
(multistep history, newest first: M X^{n-i} and F(X^{n-i})).  This module writes the same keys
in the reference's NumPy format (``.npz``, written to a temporary file and renamed, so that a job
killed while writing never leaves a truncated checkpoint); the state is stored in this package's (K, n)
mode layout.  The HDF5 variant needs h5py, which is not installed: ``.h5`` paths are refused.

Restart compares the numerics (scheme, dt, pivoting) and the identity of the problem (grid,
walls, dealiasing, Ra, Pr) and refuses a mismatch; the arrays are restored without conversion,
so the continued run is bit-identical to the uninterrupted one (tests/test_gpu_checkpoint.py).
Servers use the in-memory payload (``payload`` / ``restore``) to evict and resume runs.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from .errors import OptionError

__all__ = ["FORMAT", "payload", "restore", "save_checkpoint", "load_checkpoint"]

FORMAT = "conatus.checkpoint/1"


def _identity(pb) -> dict:
    return pb.identity()


def payload(solver) -> tuple[dict, dict]:
    """(meta, host arrays) of an IVPSolver: state, multistep history (newest first) and clock."""
    im = solver.imex
    if getattr(im, "cn", None) is not None:
        raise OptionError(f"checkpoint: {solver.scheme} runs carry an L X history that the "
                          f"{FORMAT} keys (X, MX<i>, F<i>) do not hold",
                          fix="checkpoint SBDF or RK runs")
    meta = {"format": FORMAT, "scheme": solver.scheme, "dt": float(solver.dt), "pivot": bool(solver.pivot),
            "iteration": int(im.iteration), "t": float(solver.t), "n": int(solver.problem.n),
            "K": int(solver.problem.K), "nhist": int(im.history_length), "identity": _identity(solver.problem)}
    arrays = {"X": solver.state()}
    if im.order:
        shape = (im.batch, im.n)
        for i, slot in enumerate(im._levels):
            arrays[f"MX{i}"] = im._ring[slot][0].view(shape).cpu().numpy()
            arrays[f"F{i}"] = im._ring[slot][1].view(shape).cpu().numpy()
    return meta, arrays


def restore(solver, meta: dict, arrays: dict, where: str = "checkpoint") -> None:
    """Continue ``solver`` from a payload: numerics and problem identity must match."""
    torch = __import__("torch")
    if meta.get("format") != FORMAT:
        raise OptionError(f"{where}: not a Conatus checkpoint (format {meta.get('format')!r}, expected {FORMAT!r})")
    for key, have in (("scheme", solver.scheme), ("dt", float(solver.dt)), ("pivot", bool(solver.pivot))):
        if meta.get(key) != have:
            raise OptionError(f"{where}: written with {key} = {meta.get(key)!r}, the solver has {have!r}",
                              fix="restart with the numerics the checkpoint was written with")
    mine = _identity(solver.problem)
    if meta.get("identity") != mine:
        diff = sorted(k for k in mine if meta.get("identity", {}).get(k) != mine[k])
        raise OptionError(f"{where}: the checkpoint is a state of another problem (differs in {', '.join(diff)})")
    im = solver.imex
    X = np.asarray(arrays["X"])
    if X.shape != (solver.problem.K, solver.problem.n):
        raise OptionError(f"{where}: state of shape {X.shape}, expected {(solver.problem.K, solver.problem.n)}")
    q = int(meta.get("nhist", 0))
    if im.order and q > im.order:
        raise OptionError(f"{where}: {q} history levels for {solver.scheme}")
    solver.X.copy_(torch.from_numpy(np.ascontiguousarray(X, dtype=np.complex128)))
    solver.t = float(meta["t"])
    im.reset()
    if im.order:
        shape = (im.batch, im.n)
        for i in range(q):
            im._ring[i][0].view(shape).copy_(torch.from_numpy(np.ascontiguousarray(arrays[f"MX{i}"])))
            im._ring[i][1].view(shape).copy_(torch.from_numpy(np.ascontiguousarray(arrays[f"F{i}"])))
        im._levels = list(range(q))
    im.iteration = int(meta["iteration"])
    im._remember(solver.X, solver.t)


def save_checkpoint(solver, path: str | Path) -> Path:
    """Write the payload of ``solver`` to ``path`` (``.npz``; the suffix is added if missing)."""
    path = Path(path)
    if path.suffix in (".h5", ".hdf5"):
        raise OptionError("save_checkpoint: HDF5 checkpoints need h5py, which is not available",
                          fix="save_checkpoint('state.npz')")
    if not str(path).endswith(".npz"):
        path = Path(str(path) + ".npz")
    meta, arrays = payload(solver)
    part = path.with_name(path.name[:-4] + ".part.npz")
    np.savez(part, meta=json.dumps(meta), **arrays)
    os.replace(part, path)
    return path


def load_checkpoint(solver, path: str | Path) -> None:
    """Continue ``solver`` from a file written by ``save_checkpoint`` (pickled objects refused)."""
    path = Path(path)
    if not path.exists() and path.with_suffix(".npz").exists():
        path = path.with_suffix(".npz")
    if not path.exists():
        raise OptionError(f"load_checkpoint: {path} does not exist")
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    meta = json.loads(str(arrays.pop("meta")))
    restore(solver, meta, arrays, where=str(path))
