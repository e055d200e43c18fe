"""Fourier x Chebyshev problems of the IMEX path: 2D Rayleigh-Benard convection and the 3D
plane channel (Channel3D below).

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

    field_names = ("u", "w", "T", "p")
    tau_names = TAUS

    def unknown_name(self, j: int) -> str:
        """Name of unknown j of a mode's state vector: a field coefficient ('T[5]') or a tau."""
        for f, name in enumerate(self.field_names):
            hit = np.flatnonzero(self.col(np.arange(self.N), f) == j)
            if hit.size:
                return f"{name}[{int(hit[0])}]"
        for t in self.tau_names:
            if self.tau_col(t) == j:
                return t
        return f"unknown {j}"

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

    kind = "rb2d"

    def discretization_key(self) -> tuple:
        """What members of one ensemble must share: the grid and the dealiasing (not Ra, Pr)."""
        return (self.Nx, self.Nz, float(self.Lx), float(self.z0), float(self.z1), float(self.dealias))

    def params(self) -> tuple:
        """The parameters the implicit operators depend on."""
        return (float(self.Ra), float(self.Pr))

    def describe(self) -> str:
        return f"RB 2D {self.Nx}x{self.Nz}, Ra={self.Ra:g}"

    def explicit_scratch_bytes(self) -> int:
        return 9 * self.K * self.Mz * 16

    def identity(self) -> dict:
        """What a checkpoint of this problem must match on restart."""
        return {"problem": "rayleigh_benard_2d", "Nx": int(self.Nx), "Nz": int(self.Nz), "Lx": float(self.Lx),
                "z0": float(self.z0), "z1": float(self.z1), "dealias": float(self.dealias), "Ra": float(self.Ra),
                "Pr": float(self.Pr)}

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


# ------------------------------------------------------------------------------------------------
# 3D plane channel
# ------------------------------------------------------------------------------------------------

UX, UY, UZ = 0, 1, 2                       # velocity components at column 1 + 4 n + f (P = 3)
CH_TAUS = ("tu2x", "tu2y", "tu2z", "tu1x", "tu1y", "tu1z", "tp")


@dataclass
class Channel3D:
    """Incompressible Navier-Stokes between walls y = -1, 1, periodic in x (length Lx) and z
    (length Lz), primitive variables (u, p), driven by a constant pressure gradient G e_x
    (tests/channel3d/ch3d_cases.py:20-78, the reference's channel case with
    ``pressure_gradient=G``):

        div(u) + ey . lift(tu2, -1) + tp = 0
        dt(u) - lap(u)/Re + grad(p) + lift(tu1, -1) + lift(tu2, -2) = -u . grad(u) + G ex
        u(-1) = u(1) = 0, integ(p) = 0   (mode (0, 0) gauge: tp)

    G = 2/Re makes the laminar Poiseuille flow u = (1 - y^2) ex a steady state.

    Discretization: one complex almost-banded system per Fourier mode exp(i(kx kx0 x + kz kz0 z)),
    kx = 0..Kx-1 (Kx = Nx/2) and kz = 0..Nz/2-1, -(Nz/2-1)..-1 (Kz = Nz - 1 values in FFT order,
    the Nyquist modes not represented); mode index m = kx * Kz + iz.  The fields are real, so the
    kx = 0 modes hold conjugate pairs (kz, -kz).  Within a mode the layout is the one of
    RayleighBenard2D with the wall-normal velocity v in the slot of w and the spanwise velocity w
    in the slot of T: unknowns and equation rows interleaved by Chebyshev index, equations in C^(2)
    (momentum) and C^(1) (continuity), taus lifted into T_{N-1}, T_{N-2}, border rows w(+-1),
    u(+-1), v(+-1) and the gauge, pinned lowest modes u0, u1, w0, w1, v0, p0 in the first columns.
    Every mode has n = 4 Ny + 7 and r = 7.
    """
    Nx: int
    Ny: int
    Nz: int
    Re: float
    Lx: float = 4 * np.pi
    Lz: float = 2 * np.pi
    G: float | None = None
    dealias: float = 1.5
    mats: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if self.Nx % 2 or self.Nz % 2:
            raise ValueError("Nx and Nz must be even (the Nyquist modes are not represented)")
        if self.G is None:
            self.G = 2.0 / self.Re
        self.Kx, self.Kz = self.Nx // 2, self.Nz - 1
        self.K = self.Kx * self.Kz
        self.N = self.Ny
        self.kx0 = 2 * np.pi / self.Lx
        self.kz0 = 2 * np.pi / self.Lz
        self.Mx = S.dealias_size(self.Nx, self.dealias)
        self.My = S.dealias_size(self.Ny, self.dealias)
        self.Mz = S.dealias_size(self.Nz, self.dealias)
        self.n = 4 * self.N + 7
        self.border = list(range(NBORDER))
        N = self.N
        self.S01 = S.conversion(N, 0)
        self.S12 = S.conversion(N, 1)
        self.S02 = S.conversion_chain(N, 0, 2)
        self.D1 = S.differentiation(N, 0, 1)
        self.D2 = S.differentiation(N, 0, 2)
        self._terms()
        self.mats = [self.mode_matrices(m) for m in range(self.K)]

    # -- modes and layout --------------------------------------------------------------------------
    def kz_of(self, iz):
        iz = np.asarray(iz)
        return np.where(iz < self.Nz // 2, iz, iz - self.Kz)

    def mode(self, kx: int, kz: int) -> int:
        return int(kx * self.Kz + (kz if kz >= 0 else kz + self.Kz))

    def wavenumbers(self, m: int) -> tuple[int, int]:
        return int(m // self.Kz), int(self.kz_of(m % self.Kz))

    col = RayleighBenard2D.col
    row = RayleighBenard2D.row
    field_index = RayleighBenard2D.field_index
    equation_rows = RayleighBenard2D.equation_rows
    unknown_name = RayleighBenard2D.unknown_name
    field_names = ("u", "v", "w", "p")
    tau_names = CH_TAUS

    def tau_col(self, name):
        if name == "tp":
            return 0
        return 1 + 4 * self.N + CH_TAUS.index(name)

    # -- operators ---------------------------------------------------------------------------------
    def _csr(self, R, C, V):
        A = sp.csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(self.n, self.n))
        A.sum_duplicates()
        return A

    def _put(self, R, C, V, block, e, f):
        B = sp.coo_matrix(block)
        R.append(self.row(B.row, e))
        C.append(self.col(B.col, f))
        V.append(B.data.astype(complex))

    def _put_col(self, R, C, V, vec, e, col):
        nz = np.flatnonzero(vec)
        R.append(self.row(nz, e))
        C.append(np.full(nz.size, col))
        V.append(np.asarray(vec[nz], dtype=complex))

    def _terms(self):
        """The mode-independent parts of L and the coefficients of kx^2 + kz^2, i kx and i kz."""
        N = self.N
        S02, S01, ReInv = self.S02, self.S01, 1.0 / self.Re
        R, C, V = [], [], []
        for e, f in ((XMOM, UX), (ZMOM, UY), (TEQ, UZ)):
            self._put(R, C, V, S02, e, f)
        self.M = self._csr(R, C, V)
        # L0: -D2/Re, the wall-normal pressure gradient, continuity dv/dy, lifts, wall rows
        R, C, V = [], [], []
        for e, f in ((XMOM, UX), (ZMOM, UY), (TEQ, UZ)):
            self._put(R, C, V, -ReInv * self.D2, e, f)
        self._put(R, C, V, (self.S12 @ self.D1).tocsr(), ZMOM, P)
        self._put(R, C, V, self.D1, CONT, UY)
        S02d, S01d = S02.toarray(), S01.toarray()
        for e, c in ((XMOM, "x"), (ZMOM, "y"), (TEQ, "z")):
            self._put_col(R, C, V, S02d[:, N - 1], e, self.tau_col("tu1" + c))
            self._put_col(R, C, V, S02d[:, N - 2], e, self.tau_col("tu2" + c))
        self._put_col(R, C, V, S01d[:, N - 1], CONT, self.tau_col("tu2y"))
        left, right = (-1.0) ** np.arange(N), np.ones(N)
        for r, (vals, f) in enumerate(((left, UZ), (right, UZ), (left, UX), (right, UX), (left, UY), (right, UY))):
            R.append(np.full(N, r))
            C.append(self.col(np.arange(N), f))
            V.append(vals.astype(complex))
        self.L0 = self._csr(R, C, V)
        # gauge row: integ(p) = 0 and the tp column of continuity at (0, 0); tp = 0 elsewhere
        R, C, V = [], [], []
        w = S.integration_weights(N, 2.0)
        nz = np.flatnonzero(w)
        R.append(np.full(nz.size, 6))
        C.append(self.col(nz, P))
        V.append(w[nz].astype(complex))
        self._put_col(R, C, V, S01d[:, 0], CONT, self.tau_col("tp"))
        self.Lg0 = self._csr(R, C, V)
        self.Lg = self._csr([np.array([6])], [np.array([self.tau_col("tp")])], [np.array([1.0 + 0j])])
        # k^2/Re S02 on the momentum rows; i kx: x pressure gradient + du/dx; i kz: z gradient + dw/dz
        R, C, V = [], [], []
        for e, f in ((XMOM, UX), (ZMOM, UY), (TEQ, UZ)):
            self._put(R, C, V, ReInv * S02, e, f)
        self.Lk2 = self._csr(R, C, V)
        R, C, V = [], [], []
        self._put(R, C, V, S02, XMOM, P)
        self._put(R, C, V, S01, CONT, UX)
        self.Lkx = self._csr(R, C, V)
        R, C, V = [], [], []
        self._put(R, C, V, S02, TEQ, P)
        self._put(R, C, V, S01, CONT, UZ)
        self.Lkz = self._csr(R, C, V)

    def mode_matrices(self, m: int):
        """(M, L) of mode m (complex CSR of size n; the border rows are rows 0..6):
        L = L0 + (kx^2 + kz^2)/Re S02 + i kx (S02 p_x-row, S01 u-continuity) + i kz (...)."""
        kx, kz = self.wavenumbers(m)
        ax, az = kx * self.kx0, kz * self.kz0
        L = self.L0 + (self.Lg0 if (kx, kz) == (0, 0) else self.Lg)
        if kx or kz:
            L = L + (ax * ax + az * az) * self.Lk2
        if kx:
            L = L + (1j * ax) * self.Lkx
        if kz:
            L = L + (1j * az) * self.Lkz
        L = L.tocsr()
        L.sum_duplicates()
        L.eliminate_zeros()
        return self.M, L

    def b0(self) -> np.ndarray:
        """(K, n) constant part of F: the driving pressure gradient G (T_0 coefficient of the x
        momentum, converted to C^(2)) at mode (0, 0); the wall rows are homogeneous."""
        b = np.zeros((self.K, self.n), complex)
        b[0, self.equation_rows(XMOM)] = self.G * self.S02.toarray()[:, 0]
        return b

    kind = "channel3d"

    def discretization_key(self) -> tuple:
        """What members of one ensemble must share: the grid, the dealiasing and the driving
        pressure gradient (the constant part of F); Re may differ."""
        return ("channel3d", self.Nx, self.Ny, self.Nz, float(self.Lx), float(self.Lz), float(self.dealias),
                float(self.G))

    def params(self) -> tuple:
        """The parameters the implicit operators depend on."""
        return (float(self.Re),)

    def describe(self) -> str:
        return f"channel 3D {self.Nx}x{self.Ny}x{self.Nz}, Re={self.Re:g}"

    def explicit_scratch_bytes(self) -> int:
        """Workspace of the explicit term (the intermediate spectra between its launches)."""
        Kx, Kz, Mx, My, Mz = self.Kx, self.Kz, self.Mx, self.My, self.Mz
        r1 = max(Kx * My * 6 * Kz, My * Mz * 3 * Kx)
        r2 = max(My * Mz * 9 * Kx, My * Kx * 3 * Kz)
        return 16 * (r1 + r2)

    def identity(self) -> dict:
        """What a checkpoint of this problem must match on restart."""
        return {"problem": "channel_3d", "Nx": int(self.Nx), "Ny": int(self.Ny), "Nz": int(self.Nz),
                "Lx": float(self.Lx), "Lz": float(self.Lz), "dealias": float(self.dealias), "Re": float(self.Re),
                "G": float(self.G)}

    # -- states ------------------------------------------------------------------------------------
    def velocity(self, X: np.ndarray) -> np.ndarray:
        """(3, Kx, Kz, N) Chebyshev coefficients of u, v, w from a (K, n) state."""
        return np.stack([X[:, self.field_index(f)].reshape(self.Kx, self.Kz, self.N) for f in (UX, UY, UZ)])

    def state_from_velocity(self, V: np.ndarray) -> np.ndarray:
        X = np.zeros((self.K, self.n), complex)
        for f in (UX, UY, UZ):
            X[:, self.field_index(f)] = V[f].reshape(self.K, self.N)
        return X

    def poiseuille(self) -> np.ndarray:
        """(K, n) laminar state u = (G Re / 2)(1 - y^2) ex (p = 0, taus 0)."""
        V = np.zeros((3, self.Kx, self.Kz, self.N), complex)
        a = self.G * self.Re / 2
        V[UX, 0, 0, 0] = a / 2                     # 1 - y^2 = T_0/2 - T_2/2
        V[UX, 0, 0, 2] = -a / 2
        return self.state_from_velocity(V)

    def initial_state(self, seed: int = 0, energy: float = 1e-4, kmax: int = 4) -> np.ndarray:
        """Laminar Poiseuille flow plus a seeded divergence-free perturbation u' = curl(psi),
        psi = (1 - y^2)^2 times random complex amplitudes on the modes |kx|, |kz| <= kmax
        ((0, 0) excluded; conjugate pairs at kx = 0), scaled to the volume-averaged kinetic
        energy ``energy`` (1/2 <|u'|^2>).  (1 - y^2)^2 vanishes with its y-derivative at the
        walls, so u' satisfies no-slip; the curl is taken on the coefficients, so div u' = 0."""
        rng = np.random.default_rng(seed)
        Kx, Kz, N = self.Kx, self.Kz, self.N
        prof = np.zeros(N)
        prof[:5] = np.array([3.0, 0.0, -4.0, 0.0, 1.0]) / 8.0          # (1 - y^2)^2 in T_n
        psi = np.zeros((3, Kx, Kz, N), complex)
        for kx in range(min(kmax, Kx - 1) + 1):
            for kz in range(-min(kmax, self.Nz // 2 - 1), min(kmax, self.Nz // 2 - 1) + 1):
                if (kx, kz) == (0, 0) or (kx == 0 and kz < 0):
                    continue
                a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
                psi[:, kx, self.mode(kx, kz) % Kz] = a[:, None] * prof
                if kx == 0:
                    psi[:, 0, self.mode(0, -kz) % Kz] = np.conj(a)[:, None] * prof
        ikx = (1j * self.kx0 * np.arange(Kx))[:, None, None]
        ikz = (1j * self.kz0 * self.kz_of(np.arange(Kz)))[None, :, None]
        dy = lambda c: _cheb_diff_last(c)                              # noqa: E731
        up = np.stack([dy(psi[2]) - ikz * psi[1], ikz * psi[0] - ikx * psi[2], ikx * psi[1] - dy(psi[0])])
        e = self.kinetic_energy(up)
        if e > 0:
            up *= np.sqrt(energy / e)
        return self.poiseuille() + self.state_from_velocity(up)

    def kinetic_energy(self, V: np.ndarray) -> float:
        """1/2 volume average of |u|^2 from (3, Kx, Kz, N) coefficients (Parseval in x, z,
        Gauss-Chebyshev quadrature of the exact polynomial product in y)."""
        My = 2 * self.N
        B = S.cheb_backward_matrix(self.N, My)
        vals = np.einsum("jn,fxzn->fxzj", B, V)
        w = np.full(self.Kx, 2.0)
        w[0] = 1.0
        t = S.cheb_nodes(My)
        q = np.pi / My * np.sqrt(1 - t * t)                             # Gauss-Chebyshev weights of dy
        return 0.5 * float(np.einsum("x,fxzj,j->", w, np.abs(vals) ** 2, q)) / 2.0


def _cheb_diff_last(c: np.ndarray) -> np.ndarray:
    """d/dy on [-1, 1] of Chebyshev series along the last axis (backward recurrence)."""
    N = c.shape[-1]
    out = np.zeros_like(c)
    if N > 1:
        out[..., N - 2] = 2 * (N - 1) * c[..., N - 1]
    for k in range(N - 3, -1, -1):
        out[..., k] = out[..., k + 2] + 2 * (k + 1) * c[..., k + 1]
    out[..., 0] *= 0.5
    return out
