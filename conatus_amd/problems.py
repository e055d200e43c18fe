"""Fourier x Chebyshev problems of the IMEX path: 2D Rayleigh-Benard convection.

Equations (python/conatus/problems/convection.py:24-63, tests/ivp/ivp_cases.py:59-93), x periodic
of length Lx, z in [z0, z1] between rigid walls, Boussinesq in thermal-diffusion units:

    div(u) + ez . lift(tu2, -1) + tp = 0
    dt(T) - lap(T) + lift(tT1, -1) + lift(tT2, -2) = -u . grad(T)
    dt(u) - Pr lap(u) + grad(p) - Pr Ra T ez + lift(tu1, -1) + lift(tu2, -2) = -u . grad(u)
    T(z0) = +1/2, T(z1) = -1/2, u(z0) = u(z1) = 0, integ(p) = 0   (k = 0 gauge: tp)

Discretization (the reference's conventions): for every complex Fourier mode k (wavenumber
k kappa, kappa = 2 pi / Lx) an almost-banded system of Chebyshev coefficients, equations in the
ultraspherical basis C^(lam) of their highest z-derivative (lam = 2 for momentum and
temperature, 1 for continuity), taus lifted into T_{N-1} / T_{N-2} and converted with the
equation (the generalized tau of the continuity row), the gauge tp = the T_0 coefficient of the
continuity equation at k = 0.  Unknowns and equation rows are interleaved by Chebyshev index
(u, w, T, p at index n next to each other), which makes the bulk rows banded; the boundary rows
are the border.  The first r columns hold what only the border rows determine (the reference's
"pinned lowest modes", python/conatus/timestep/system.py:418-533): the gauge tp and u0, u1, T0,
T1, w0, p0, so that the banded part B of the Woodbury split is invertible for every mode; the six
lifted taus follow the last Chebyshev index.  Every mode has the same size
n = 4 Nz + 7: the modes k > 0 carry the gauge slot as a trivial row tp = 0, so that one batch holds
all modes.

The explicit term F of the momentum and temperature rows is S_{0->2} applied to the Chebyshev
coefficients of -u.grad(u) and -u.grad(T), computed on the 3/2-dealiased grid; the boundary
rows of F carry the constant right-hand sides (T = +-1/2 at k = 0).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import spectral as S

# field and equation slots of the interleaved layout
U, W, TT, P = 0, 1, 2, 3                  # unknown fields at column 1 + 4 n + f (column 0: tp)
XMOM, ZMOM, TEQ, CONT = 0, 1, 2, 3        # equation rows at R + 4 n + e
TAUS = ("tu2x", "tu2z", "tT2", "tu1x", "tu1z", "tT1", "tp")   # lift(-2) taus reach further up: first
NBORDER = 7                               # T(z0), T(z1), u(z0), u(z1), w(z0), w(z1), gauge


@dataclass
class RayleighBenard2D:
    Nx: int
    Nz: int
    Lx: float
    Ra: float
    Pr: float
    z0: float = -1.0
    z1: float = 1.0
    dealias: float = 1.5
    mats: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if self.Nx % 2:
            raise ValueError("Nx must be even (the Nyquist mode is not represented)")
        self.K = self.Nx // 2
        self.N = self.Nz
        self.Lz = self.z1 - self.z0
        self.kappa = 2 * np.pi / self.Lx
        self.Mx = S.dealias_size(self.Nx, self.dealias)
        self.Mz = S.dealias_size(self.Nz, self.dealias)
        self.n = 4 * self.N + 7
        self.border = list(range(NBORDER))
        N = self.N
        self.S01 = S.conversion(N, 0)
        self.S12 = S.conversion(N, 1)
        self.S02 = S.conversion_chain(N, 0, 2)
        self.D1 = S.differentiation(N, 0, 1) * (2.0 / self.Lz)
        self.D2 = S.differentiation(N, 0, 2) * (2.0 / self.Lz) ** 2
        self.mats = [self.mode_matrices(k) for k in range(self.K)]

    # -- layout ------------------------------------------------------------------------------------
    def col(self, n, f):
        """Column of coefficient n of field f: the pinned lowest modes (u0, u1, T0, T1, w0, p0, which
        only the boundary rows determine) sit in columns 1..6 after tp, then w1, p1, then the
        interleaved block 1 + 4 n + f for n >= 2."""
        n = np.asarray(n)
        c = 1 + 4 * n + f
        low = {(0, U): 1, (1, U): 2, (0, TT): 3, (1, TT): 4, (0, W): 5, (0, P): 6, (1, W): 7, (1, P): 8}
        out = np.array(c, copy=True)
        for (m, ff), v in low.items():
            if ff == f:
                out = np.where(n == m, v, out)
        return out if out.ndim else int(out)

    def row(self, n, e):
        return NBORDER + 4 * np.asarray(n) + e

    def tau_col(self, name):
        if name == "tp":
            return 0
        return 1 + 4 * self.N + TAUS.index(name)

    def field_index(self, f) -> np.ndarray:
        """Positions of field f's N Chebyshev coefficients in a mode's state vector."""
        return self.col(np.arange(self.N), f)

    def equation_rows(self, e) -> np.ndarray:
        return NBORDER + 4 * np.arange(self.N) + e

    # -- operators ---------------------------------------------------------------------------------
    def mode_matrices(self, k: int):
        """(M, L) of mode k, complex CSR of size n; the border rows are rows 0..6."""
        N, n = self.N, self.n
        kk = k * self.kappa
        Mr, Mc, Mv = [], [], []
        Lr, Lc, Lv = [], [], []

        def put(R, C, V, block, e, f):
            B = sp.coo_matrix(block)
            R.append(self.row(B.row, e))
            C.append(self.col(B.col, f))
            V.append(B.data.astype(complex))

        def put_col(R, C, V, vec, e, col):
            nz = np.flatnonzero(vec)
            R.append(self.row(nz, e))
            C.append(np.full(nz.size, col))
            V.append(np.asarray(vec[nz], dtype=complex))

        lap2 = (self.D2 - kk ** 2 * self.S02).tocsr()
        S02, S01 = self.S02, self.S01
        # mass: dt(u), dt(w), dt(T) in C^(2)
        for e, f in ((XMOM, U), (ZMOM, W), (TEQ, TT)):
            put(Mr, Mc, Mv, S02, e, f)
        # x momentum
        put(Lr, Lc, Lv, -self.Pr * lap2, XMOM, U)
        if k:
            put(Lr, Lc, Lv, 1j * kk * S02, XMOM, P)
        # z momentum
        put(Lr, Lc, Lv, -self.Pr * lap2, ZMOM, W)
        put(Lr, Lc, Lv, (self.S12 @ self.D1).tocsr(), ZMOM, P)
        put(Lr, Lc, Lv, -self.Pr * self.Ra * S02, ZMOM, TT)
        # temperature
        put(Lr, Lc, Lv, -lap2, TEQ, TT)
        # continuity (C^(1))
        if k:
            put(Lr, Lc, Lv, 1j * kk * S01, CONT, U)
        put(Lr, Lc, Lv, self.D1, CONT, W)
        # lifts into T_{N-1}, T_{N-2}, converted with the equation
        S02d = S02.toarray()
        S01d = S01.toarray()
        put_col(Lr, Lc, Lv, S02d[:, N - 1], XMOM, self.tau_col("tu1x"))
        put_col(Lr, Lc, Lv, S02d[:, N - 2], XMOM, self.tau_col("tu2x"))
        put_col(Lr, Lc, Lv, S02d[:, N - 1], ZMOM, self.tau_col("tu1z"))
        put_col(Lr, Lc, Lv, S02d[:, N - 2], ZMOM, self.tau_col("tu2z"))
        put_col(Lr, Lc, Lv, S02d[:, N - 1], TEQ, self.tau_col("tT1"))
        put_col(Lr, Lc, Lv, S02d[:, N - 2], TEQ, self.tau_col("tT2"))
        put_col(Lr, Lc, Lv, S01d[:, N - 1], CONT, self.tau_col("tu2z"))
        if k == 0:
            put_col(Lr, Lc, Lv, S01d[:, 0], CONT, self.tau_col("tp"))
        # boundary rows (interp at z0: T_n(-1) = (-1)^n; at z1: T_n(1) = 1)
        left = (-1.0) ** np.arange(N)
        right = np.ones(N)
        for r, (vals, f) in enumerate(((left, TT), (right, TT), (left, U), (right, U), (left, W), (right, W))):
            Lr.append(np.full(N, r))
            Lc.append(self.col(np.arange(N), f))
            Lv.append(vals.astype(complex))
        if k == 0:
            w = S.integration_weights(N, self.Lz)
            nz = np.flatnonzero(w)
            Lr.append(np.full(nz.size, 6))
            Lc.append(self.col(nz, P))
            Lv.append(w[nz].astype(complex))
        else:
            Lr.append(np.array([6]))
            Lc.append(np.array([self.tau_col("tp")]))
            Lv.append(np.array([1.0 + 0j]))
        M = sp.csr_matrix((np.concatenate(Mv), (np.concatenate(Mr), np.concatenate(Mc))), shape=(n, n))
        L = sp.csr_matrix((np.concatenate(Lv), (np.concatenate(Lr), np.concatenate(Lc))), shape=(n, n))
        M.sum_duplicates()
        L.sum_duplicates()
        return M, L

    def b0(self) -> np.ndarray:
        """(K, n) constant right-hand side of the boundary rows (T = +1/2 at z0, -1/2 at z1)."""
        b = np.zeros((self.K, self.n), complex)
        b[0, 0] = 0.5
        b[0, 1] = -0.5
        return b

    # -- grids and states --------------------------------------------------------------------------
    def grids(self, dealias: bool = True):
        Mx, Mz = (self.Mx, self.Mz) if dealias else (self.Nx, self.Nz)
        return S.fourier_grid(Mx, 0.0, self.Lx), S.cheb_grid(Mz, self.z0, self.z1)

    def coeffs_from_grid(self, f: np.ndarray) -> np.ndarray:
        """(Mx, Mz) real grid values -> (K, N) complex coefficients (forward x then z)."""
        Fx = S.fourier_forward_matrix(self.K, f.shape[0]) @ f          # (K, Mz)
        return Fx @ S.cheb_forward_matrix(self.N, f.shape[1]).T        # (K, N)

    def grid_from_coeffs(self, c: np.ndarray, Mx=None, Mz=None) -> np.ndarray:
        Mx = Mx or self.Mx
        Mz = Mz or self.Mz
        cz = c @ S.cheb_backward_matrix(self.N, Mz).T                  # (K, Mz)
        return np.real(S.fourier_backward_matrix(self.K, Mx, 0, self.Lx) @ cz)

    def state_from_fields(self, T=None, u=None, w=None) -> np.ndarray:
        """(K, n) state from grid fields on the dealiased grid (p and taus zero)."""
        X = np.zeros((self.K, self.n), complex)
        for f, g in ((TT, T), (U, u), (W, w)):
            if g is not None:
                X[:, self.field_index(f)] = self.coeffs_from_grid(np.asarray(g, float))
        return X

    def initial_state(self, seed: int = 0, amplitude: float = 1e-3) -> np.ndarray:
        """Conduction profile T = 1/2 - (z - z0)/Lz plus seeded Gaussian grid noise of the given
        amplitude times the wall envelope (1 - t^2), on the dealiased grid; u = w = 0."""
        x, z = self.grids()
        t = 2 * (z - self.z0) / self.Lz - 1
        rng = np.random.default_rng(seed)
        T = 0.5 - (z[None, :] - self.z0) / self.Lz + amplitude * rng.standard_normal((x.size, z.size)) * (1 - t[None, :] ** 2)
        return self.state_from_fields(T=T)
