"""RB 2D discretization against the reference's own (tests/golden/rb2d_modes_8x8.npz, written by
tests/golden/make_rb2d_golden.py from the reference's rayleigh_benard IVP,
python/conatus/problems/convection.py:26-72, discretized by its discretize + GlobalSystem).

The two layouts differ (the reference stacks each kx >= 1 mode as a real system of cos / msin
coefficient pairs; conatus_amd keeps one complex system per mode, border rows first), so the pin
is layout-free: the finite generalized eigenvalues of every mode's pencil (sigma M + L) x = 0 --
the decay and growth rates of the discrete linear operator, which a tau row in another place, a
boundary row of another kind or a different lift would change -- agree to 1e-10 relative on the
reference's domain z in [-1/2, 1/2] (each eigenvalue of a real cos / msin system appears twice)."""
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sl

GOLDEN = Path(__file__).resolve().parent / "golden" / "rb2d_modes_8x8.npz"


def _finite(ev):
    ev = ev[np.isfinite(ev)]
    return np.sort_complex(ev[np.abs(ev) < 1e8])


@pytest.mark.skipif(not GOLDEN.exists(), reason="golden fixture not generated")
def test_mode_pencils_have_the_references_spectra():
    from conatus_amd.problems import RayleighBenard2D
    g = np.load(GOLDEN, allow_pickle=False)
    pb = RayleighBenard2D(Nx=int(g["Nx"]), Nz=int(g["Nz"]), Lx=float(g["Lx"]), Ra=float(g["Ra"]),
                          Pr=float(g["Pr"]), z0=-0.5, z1=0.5)
    assert int(g["slots"]) == pb.K
    for k in range(pb.K):
        ref = _finite(sl.eig(-g[f"L{k}"], g[f"M{k}"], right=False))
        if k > 0:
            assert np.max(abs(ref[0::2] - ref[1::2])) <= 1e-9 * np.max(abs(ref))   # cos / msin pairs
            ref = ref[0::2]
        M, L = pb.mats[k]
        ours = _finite(sl.eig(-L.toarray(), M.toarray(), right=False))
        assert len(ours) == len(ref), (k, len(ours), len(ref))
        assert np.max(abs(ours - ref)) <= 1e-10 * np.max(abs(ref)), k
