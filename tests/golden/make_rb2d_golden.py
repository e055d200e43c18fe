"""Writes tests/golden/rb2d_modes_8x8.npz: the reference's own per-mode matrices (M, L), border
rows and unknown ordering of its 2D Rayleigh-Benard IVP at a tiny grid, so that the tau placement
and row/column layout of conatus_amd.problems.RayleighBenard2D are pinned to the reference's
discretization rather than only to the physics.

Run in the build container only (it reads /root/reference, which the GPU box does not have):

    python tests/golden/make_rb2d_golden.py

The reference's runtime loads its compiled core (conatus._core) at import; that core is not built
here, so the stand-in conatus_amd.core_shim (the same binding names) is registered under that name
first.  Only the discretization is used (python/conatus/problems/convection.py:26-72,
discretize/plan.py:1111, timestep/system.py:244-330); nothing is solved.
"""
from __future__ import annotations

import importlib.machinery
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
REFERENCE_PY = Path("/root/reference/python")


def main() -> int:
    sys.path.insert(0, str(REPO))
    sys.path.insert(0, str(REFERENCE_PY))
    from conatus_amd import core_shim
    core_shim.__spec__ = importlib.machinery.ModuleSpec("conatus._core", None)
    sys.modules["conatus._core"] = core_shim
    from conatus.discretize.plan import discretize
    from conatus.problems.convection import rayleigh_benard
    from conatus.timestep.system import GlobalSystem

    Nx, Nz, Lx, Ra, Pr = 8, 8, 4.0, 2500.0, 1.0
    rb = rayleigh_benard(Nx=Nx, Nz=Nz, Lx=Lx, Ra=Ra, Pr=Pr, dealias=1.5)
    gs = GlobalSystem(discretize(rb.problem))
    out = {"Nx": Nx, "Nz": Nz, "Lx": Lx, "Ra": Ra, "Pr": Pr, "slots": len(gs.slots),
           "variables": np.array([v.name for v in gs.variables])}
    for i, s in enumerate(gs.slots):
        M, L = s.M.toarray(), s.L.toarray()
        out[f"M{i}"] = M
        out[f"L{i}"] = L
        out[f"key{i}"] = np.array(repr(getattr(s, "key", i)))
    np.savez_compressed(HERE / "rb2d_modes_8x8.npz", **out)
    print(f"wrote {len(gs.slots)} slots of sizes {[s.M.shape[0] for s in gs.slots]}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
