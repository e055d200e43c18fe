"""CPU oracle of the 3D plane-channel IMEX step — TEST INFRASTRUCTURE ONLY (see oracle/banded.py).

Consumes the per-mode operators (M_m, L_m), border rows and constant part b0 built by
conatus_amd.problems.Channel3D (the discretization is data here; it is pinned separately by the
reference's manufactured Stokes-resolvent mode, tests/channel3d/stokes_mode.py and
tests/channel3d/test_influence.py:50-63: error 6.8e-11 on the N = 16 grid, < 1e-13 at N = 24,
and by the laminar Poiseuille steady state).  The explicit term is restated in numpy with FFTs:

* the coefficients of u, v, w per mode (kx, kz); x- and z-derivatives by i kx kx0, i kz kz0 on
  the modes, the y-derivative by the Chebyshev backward recurrence (oracle.spectral.cheb_diff,
  python/conatus/timestep/spectral.py:53-82);
* backward transforms to the 3/2 grid: Chebyshev (oracle.spectral.cheb_backward), complex FFT in
  z over the modes kz = -(Nz/2-1)..Nz/2-1 zero-padded to Mz, real inverse FFT in x (the kx = 0
  plane's imaginary part dropped: real fields), as the reference's RealFourier / Chebyshev
  transforms (python/conatus/reference/transforms.py:103-177, timestep/xp_transforms.py:197-548);
* the products -u.grad(u) on the grid; forward transforms (rfft in x, fft in z, Chebyshev
  forward; truncation to the represented modes = the dealiasing projection), conversion S_{0->2}
  into the three momentum rows, plus the constant b0 (the driving pressure gradient).

Stepping (SBDF, IMEX RK) is RB2DOracle's, per mode, with LAPACK solves (oracle.banded).
"""
from __future__ import annotations

import numpy as np

from . import spectral as S
from .rb2d import RB2DOracle


class Channel3DOracle(RB2DOracle):
    def _zfull(self, A):
        """(Kx, Kz, ...) modes in FFT order -> (Kx, Mz, ...) zero-padded complex z spectrum."""
        pb = self.pb
        Z = np.zeros((pb.Kx, pb.Mz) + A.shape[2:], complex)
        Z[:, pb.kz_of(np.arange(pb.Kz)) % pb.Mz] = A
        return Z

    def grid(self, A: np.ndarray) -> np.ndarray:
        """(Kx, Kz, N) complex coefficients -> (Mx, Mz, My) real values on the 3/2 grid."""
        pb = self.pb
        g = S.cheb_backward(A, pb.My, axis=2)                          # (Kx, Kz, My)
        Z = np.fft.ifft(self._zfull(g), axis=1) * pb.Mz               # (Kx, Mz, My)
        Xf = np.zeros((pb.Mx // 2 + 1, pb.Mz, pb.My), complex)
        Xf[:pb.Kx] = Z
        return np.fft.irfft(Xf, n=pb.Mx, axis=0) * pb.Mx

    def modes(self, g: np.ndarray) -> np.ndarray:
        """(Mx, Mz, My) real values -> (Kx, Kz, N) coefficients (truncated transforms)."""
        pb = self.pb
        Xf = np.fft.rfft(g, axis=0)[:pb.Kx] / pb.Mx                    # (Kx, Mz, My)
        Zf = np.fft.fft(Xf, axis=1) / pb.Mz
        A = Zf[:, pb.kz_of(np.arange(pb.Kz)) % pb.Mz]                 # (Kx, Kz, My)
        return S.cheb_forward(A, pb.N, axis=2)

    def explicit(self, X: np.ndarray) -> np.ndarray:
        pb = self.pb
        V = pb.velocity(X)                                             # (3, Kx, Kz, N)
        ikx = (1j * pb.kx0 * np.arange(pb.Kx))[:, None, None]
        ikz = (1j * pb.kz0 * pb.kz_of(np.arange(pb.Kz)))[None, :, None]
        vel = [self.grid(V[f]) for f in range(3)]
        grads = [[self.grid(ikx * V[f]), self.grid(S.cheb_diff(V[f], 2.0, axis=2)), self.grid(ikz * V[f])]
                 for f in range(3)]
        F = self.b0.copy()
        for e in range(3):
            g = -(vel[0] * grads[e][0] + vel[1] * grads[e][1] + vel[2] * grads[e][2])
            c = self.modes(g).reshape(pb.K, pb.N)
            F[:, pb.equation_rows(e)] += (pb.S02 @ c.T).T
        return F
