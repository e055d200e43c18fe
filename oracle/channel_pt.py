"""CPU oracle of the poloidal-toroidal 3D channel IMEX step — TEST INFRASTRUCTURE ONLY (see
oracle/banded.py).

Restates the reference's channel_3d_pt (tests/channel3d/ch3d_cases.py:90-217) independently of
conatus_amd: its own operators from the ultraspherical closed forms of oracle.spectral, its own
explicit term with numpy FFTs, its own SBDF stepping with LAPACK solves (oracle.banded).

* Unknowns per mode (kx, kz) != (0, 0): phi (4th order in y, taus lifted into the last four
  rows of the C^(4) equation), psi (2nd order, taus in C^(2)); the mean flow U, W (2nd order);
  at (0, 0) phi and psi are pinned to zero by <phi>, <psi> (ch3d_cases.py:196-199).
  Equations, with k^2 = (kx kx0)^2 + (kz kz0)^2 and lap = D^2 - k^2:
      -k^2 [lap dt phi - lap^2 phi / Re] = D(i kx Nx + i kz Nz) + k^2 Ny          (C^(4))
      -k^2 [dt psi - lap psi / Re]       = i kz Nx - i kx Nz                     (C^(2))
      dt U - D^2 U / Re = <Nx> + G,   dt W - D^2 W / Re = <Nz>                    (C^(2))
  phi = D phi = psi = U = W = 0 at y = -1, 1.
* Velocity (ch3d_cases.py:172-174): u = i kx D phi + i kz psi (+ U), v = k^2 phi,
  w = i kz D phi - i kx psi (+ W); D by the Chebyshev recurrence (oracle.spectral.cheb_diff).
* Explicit term (the reference's evaluator path for a derivative of products,
  python/conatus/timestep/spectral.py:94-122): the six products u_i u_j on the 3/2 grid (backward
  transforms as python/conatus/reference/transforms.py:103-177: Chebyshev, complex FFT in z over
  the represented modes, real inverse FFT in x), forward transforms truncated to the represented
  modes, N_i = -(i kx P_ix + D P_iy + i kz P_iz), the right-hand sides above in the T basis, then
  converted to the equations' C^(lambda) bases.  The kx = 0 plane of a real field holds
  conjugate pairs: F(0, -kz) is taken as the conjugate of F(0, kz) and the mean rows as real.
* Stepping: SBDF1..4 as core/include/conatus/imex.hpp:158-176 on every system (the four real
  columns of a slot share the real factorization of a0/dt M + L).

State layout (the documented one of the device solver): float64 [phi: 4][P][N + 4] then
[psi: 4][P][N + 2], slot s = kx Kzh + |kz|, columns (Re, Im) of (kx, +|kz|) then of (kx, -|kz|);
phi rows [phi(-1), phi(1), phi'(-1), phi'(1), C^(4) rows], columns [phi_n, tau_4 .. tau_1];
psi rows [psi(-1), psi(1), C^(2) rows], columns [psi_n, tau_2, tau_1]; slot 0 of psi holds U
(column 0) and W (column 1).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import spectral as S
from .banded import BorderedBanded
from .imex import sbdf_coefficients


def _dense(A):
    return np.asarray(A.todense()) if sp.issparse(A) else np.asarray(A)


class ChannelPTOracle:
    def __init__(self, Nx, Ny, Nz, Re, Lx=4 * np.pi, Lz=2 * np.pi, G=None, dealias=1.5,
                 scheme: str = "SBDF3", dt: float = 5e-4):
        self.Nx, self.Ny, self.Nz, self.Re = int(Nx), int(Ny), int(Nz), float(Re)
        self.G = 2.0 / self.Re if G is None else float(G)
        self.Kx, self.Kzh, self.Kz = self.Nx // 2, self.Nz // 2, self.Nz - 1
        self.P = self.Kx * self.Kzh
        self.N = N = self.Ny
        self.kx0, self.kz0 = 2 * np.pi / Lx, 2 * np.pi / Lz
        self.Mx, self.My, self.Mz = (int(np.floor(v * dealias)) for v in (self.Nx, self.Ny, self.Nz))
        self.order = int(scheme[4])
        self.dt = float(dt)
        sc = lambda a, b: _dense(S.conversion_chain(N, a, b))            # noqa: E731
        self.C04, self.C02, self.C24 = sc(0, 4), sc(0, 2), sc(2, 4)
        self.D2 = _dense(S.differentiation(N, 0, 2))
        self.D4 = _dense(S.differentiation(N, 0, 4))
        self.MXh, self.Fh = [], []
        self.factors = {}

    # -- layout ------------------------------------------------------------------------------------
    def kz(self, iz):
        iz = np.asarray(iz)
        return np.where(iz < self.Kzh, iz, iz - self.Kz)

    def k2_slot(self, s):
        kx, kz = divmod(s, self.Kzh)
        return (kx * self.kx0) ** 2 + (kz * self.kz0) ** 2

    def flat(self, phi, psi, U, W):
        """Per-mode arrays -> the flat state (phi, psi (Kx, Kz, .) complex; U, W real)."""
        nP, nS = self.N + 4, self.N + 2
        A = np.zeros((4, self.P, nP))
        B = np.zeros((4, self.P, nS))
        for kx in range(self.Kx):
            for iz in range(self.Kz):
                k = int(self.kz(iz))
                s, c = kx * self.Kzh + abs(k), (0 if k >= 0 else 2)
                if s == 0:
                    continue
                A[c, s], A[c + 1, s] = phi[kx, iz].real, phi[kx, iz].imag
                B[c, s], B[c + 1, s] = psi[kx, iz].real, psi[kx, iz].imag
        B[0, 0], B[1, 0] = U, W
        return np.concatenate([A.ravel(), B.ravel()])

    def unflat(self, X):
        nP, nS = self.N + 4, self.N + 2
        A = X[:4 * self.P * nP].reshape(4, self.P, nP)
        B = X[4 * self.P * nP:].reshape(4, self.P, nS)
        phi = np.zeros((self.Kx, self.Kz, nP), complex)
        psi = np.zeros((self.Kx, self.Kz, nS), complex)
        for kx in range(self.Kx):
            for iz in range(self.Kz):
                k = int(self.kz(iz))
                s, c = kx * self.Kzh + abs(k), (0 if k >= 0 else 2)
                if s == 0:
                    continue
                phi[kx, iz] = A[c, s] + 1j * A[c + 1, s]
                psi[kx, iz] = B[c, s] + 1j * B[c + 1, s]
        return phi, psi, B[0, 0].copy(), B[1, 0].copy()

    # -- operators ---------------------------------------------------------------------------------
    def operators(self, bucket: str, s: int):
        """(M, L) dense real matrices of one system (rows and columns as the module doc says)."""
        N = self.N
        n = N + (4 if bucket == "phi" else 2)
        M, L = np.zeros((n, n)), np.zeros((n, n))
        nn = np.arange(N)
        left, right = (-1.0) ** nn, np.ones(N)
        k2 = self.k2_slot(s)
        if bucket == "phi":
            L[0, :N], L[1, :N] = left, right
            L[2, :N], L[3, :N] = -left * nn * nn, nn * nn
            lapC4 = self.C24 @ self.D2 - k2 * self.C04                       # lap in C^(4)
            if s == 0:
                L[4:, :N] = self.C04
            else:
                M[4:, :N] = -k2 * lapC4
                L[4:, :N] = (k2 / self.Re) * (self.D4 - 2 * k2 * (self.C24 @ self.D2) + k2 * k2 * self.C04)
            for k in range(1, 5):
                L[4 + N - k, N + 4 - k] = 1.0
        else:
            L[0, :N], L[1, :N] = left, right
            if s == 0:
                M[2:, :N] = self.C02
                L[2:, :N] = -self.D2 / self.Re
            else:
                M[2:, :N] = -k2 * self.C02
                L[2:, :N] = (k2 / self.Re) * (self.D2 - k2 * self.C02)
            for k in range(1, 3):
                L[2 + N - k, N + 2 - k] = 1.0
        return M, L

    # -- explicit term -----------------------------------------------------------------------------
    def _grid(self, A):
        """(Kx, Kz, N) complex coefficients -> (Mx, Mz, My) real values on the 3/2 grid."""
        g = S.cheb_backward(A, self.My, axis=2)
        Z = np.zeros((self.Kx, self.Mz, self.My), complex)
        Z[:, self.kz(np.arange(self.Kz)) % self.Mz] = g
        Z = np.fft.ifft(Z, axis=1) * self.Mz
        Xf = np.zeros((self.Mx // 2 + 1, self.Mz, self.My), complex)
        Xf[:self.Kx] = Z
        Xf[0] = Xf[0].real
        return np.fft.irfft(Xf, n=self.Mx, axis=0) * self.Mx

    def _modes(self, g):
        Xf = np.fft.rfft(g, axis=0)[:self.Kx] / self.Mx
        Zf = np.fft.fft(Xf, axis=1) / self.Mz
        return S.cheb_forward(Zf[:, self.kz(np.arange(self.Kz)) % self.Mz], self.N, axis=2)

    def velocity(self, phi, psi, U, W):
        N = self.N
        ax = (self.kx0 * np.arange(self.Kx))[:, None, None]
        az = (self.kz0 * self.kz(np.arange(self.Kz)))[None, :, None]
        f, g = phi[..., :N], psi[..., :N]
        df = S.cheb_diff(f, 2.0, axis=2)
        V = np.stack([1j * ax * df + 1j * az * g, (ax * ax + az * az) * f, 1j * az * df - 1j * ax * g])
        V[:, 0, 0] = 0.0
        V[0, 0, 0], V[2, 0, 0] = U[:N], W[:N]
        return V

    def explicit_modes(self, phi, psi, U, W):
        """(F_phi, F_psi, F_U, F_W) per mode, rows of the systems (boundary rows zero)."""
        N = self.N
        V = self.velocity(phi, psi, U, W)
        u, v, w = (self._grid(V[f]) for f in range(3))
        P = {key: self._modes(a * b) for key, (a, b) in
             {"uu": (u, u), "uv": (u, v), "uw": (u, w), "vv": (v, v), "vw": (v, w), "ww": (w, w)}.items()}
        ikx = (1j * self.kx0 * np.arange(self.Kx))[:, None, None]
        ikz = (1j * self.kz0 * self.kz(np.arange(self.Kz)))[None, :, None]
        k2 = -(ikx * ikx + ikz * ikz)
        D = lambda c: S.cheb_diff(c, 2.0, axis=2)                        # noqa: E731
        Nx = -(ikx * P["uu"] + D(P["uv"]) + ikz * P["uw"])
        Ny = -(ikx * P["uv"] + D(P["vv"]) + ikz * P["vw"])
        Nz = -(ikx * P["uw"] + D(P["vw"]) + ikz * P["ww"])
        rphi = D(ikx * Nx + ikz * Nz) + k2 * Ny
        rpsi = ikz * Nx - ikx * Nz
        Fphi = np.zeros((self.Kx, self.Kz, N + 4), complex)
        Fpsi = np.zeros((self.Kx, self.Kz, N + 2), complex)
        Fphi[..., 4:] = np.einsum("mn,xzn->xzm", self.C04, rphi)
        Fpsi[..., 2:] = np.einsum("mn,xzn->xzm", self.C02, rpsi)
        Fphi[0, 0] = 0.0
        Fpsi[0, 0] = 0.0
        # real fields: the kx = 0 plane is conjugate-symmetric
        for iz in range(self.Kzh, self.Kz):
            Fphi[0, iz] = np.conj(Fphi[0, self.Kz - iz])
            Fpsi[0, iz] = np.conj(Fpsi[0, self.Kz - iz])
        rU = Nx[0, 0].real.copy()
        rU[0] += self.G
        FU, FW = np.zeros(N + 2), np.zeros(N + 2)
        FU[2:] = self.C02 @ rU
        FW[2:] = self.C02 @ Nz[0, 0].real
        return Fphi, Fpsi, FU, FW

    def explicit(self, X):
        """F of a flat state, as a flat vector."""
        return self.flat(*self.explicit_modes(*self.unflat(X)))

    # -- stepping ----------------------------------------------------------------------------------
    def _systems(self, X):
        nP, nS = self.N + 4, self.N + 2
        A = X[:4 * self.P * nP].reshape(4, self.P, nP)
        B = X[4 * self.P * nP:].reshape(4, self.P, nS)
        return (("phi", A), ("psi", B))

    def mass(self, X):
        out = []
        for name, A in self._systems(X):
            Y = np.zeros_like(A)
            for s in range(self.P):
                M = sp.csr_matrix(self.operators(name, s)[0])
                for c in range(4):
                    Y[c, s] = M @ A[c, s]
            out.append(Y.ravel())
        return np.concatenate(out)

    def factor(self, name, s, alpha):
        key = (name, s, alpha)
        if key not in self.factors:
            M, L = self.operators(name, s)
            n = M.shape[0]
            self.factors[key] = BorderedBanded(sp.csr_matrix(alpha * M + L), list(range(4 if name == "phi" else 2)),
                                               np.float64)
        return self.factors[key]

    def step(self, X):
        MX = self.mass(X)
        F = self.explicit(X)
        self.MXh.insert(0, MX)
        self.Fh.insert(0, F)
        if len(self.MXh) > self.order:
            self.MXh.pop()
            self.Fh.pop()
        q = min(self.order, len(self.MXh))
        a, c = sbdf_coefficients(q)
        rhs = np.zeros_like(X)
        for j in range(1, q + 1):
            rhs = rhs + (c[j] * self.Fh[j - 1] - (a[j] / self.dt) * self.MXh[j - 1])
        out = np.empty_like(X)
        off = 0
        for name, A in self._systems(rhs):
            Y = np.empty_like(A)
            for s in range(self.P):
                Y[:, s] = self.factor(name, s, a[0] / self.dt).solve(A[:, s])
            out[off:off + Y.size] = Y.ravel()
            off += Y.size
        return out

    def run(self, X, steps):
        X = np.array(X, dtype=float)
        for _ in range(steps):
            X = self.step(X)
        return X
