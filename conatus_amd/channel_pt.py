"""3D plane channel in the poloidal-toroidal (velocity-vorticity) form.

The reference's case channel_3d_pt (tests/channel3d/ch3d_cases.py:90-217): incompressible
Navier-Stokes between walls y = -1, 1, periodic in x (length Lx) and z (length Lz), the velocity
written as u = curl curl(phi ey) + (dz psi, 0, -dx psi) + U(y) ex + W(y) ez,

    u = (dx dy phi + dz psi + U,  -lap_h phi,  dz dy phi - dx psi + W),   lap_h = dx^2 + dz^2,

divergence-free by construction.  The wall-normal curl curl and curl of the momentum equation,
with N = -div(u u) (divergence form, six products), give for every mode (kx, kz) != (0, 0)

    lap_h [dt lap phi - lap^2 phi / Re] = dy(dx Nx + dz Nz) - lap_h Ny     (order 4 in y)
    lap_h [dt psi     - lap psi / Re]   = dz Nx - dx Nz                   (order 2 in y)

with phi = dy phi = psi = 0 at the walls, and for the mean flow
dt U - U''/Re = <Nx> + G, dt W - W''/Re = <Nz>, U = W = 0 at the walls (<.>: the x-z average,
G the driving pressure gradient; G = 2/Re gives the laminar U = 1 - y^2).  Taus are lifted into
the equations' own bases: C^(4) for phi (the last four rows), C^(2) for psi, U, W (the last two).
At (0, 0) phi and psi carry no velocity; the reference pins them with <phi>, <psi>.

Discretization here (one Fourier mode exp(i(kx kx0 x + kz kz0 z)) per system, kx = 0..Kx-1,
kz = -(Nz/2-1)..Nz/2-1, the Nyquist modes not represented): every operator depends on the mode
only through k^2 = (kx kx0)^2 + (kz kz0)^2, and is real.  The modes (kx, kz) and (kx, -kz) therefore
share one real factorization, and a complex coefficient is two real right-hand sides: a slot
s = kx Kzh + |kz| (Kzh = Nz/2, P = Kx Kzh slots) holds four real columns

    c = 0: Re of mode (kx, +|kz|)   1: Im of (kx, +|kz|)   2: Re of (kx, -|kz|)   3: Im of (kx, -|kz|)

(kz = 0: columns 2 and 3 are zero).  The state is one float64 vector: the phi bucket
[4][P][N + 4] then the psi bucket [4][P][N + 2].  A phi system has the rows
[phi(-1), phi(1), phi'(-1), phi'(1), C^(4) equation rows 0..N-1] and the columns
[phi_0..phi_{N-1}, tau_4, tau_3, tau_2, tau_1] (tau_k in equation row N - k lands on the
diagonal); a psi system [psi(-1), psi(1), C^(2) rows] x [psi_n, tau_2, tau_1].  The bulk rows are
banded (phi: kl = ku = 4; psi: kl = ku = 2) and the boundary rows are the border (r = 4, 2).  Slot
0 is the mean: its phi system is the pin <phi> = 0 (S_{0->4} phi + taus, M = 0) and its psi system
holds the mean-flow operator (M = S_{0->2}, L = -D_{0->2}/Re) with U in column 0 and W in column
1.  The fields are real, so the kx = 0 modes hold conjugate pairs: column 2 = column 0 and
column 3 = -column 1 there, kept exactly by the explicit term.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import spectral as S

__all__ = ["ChannelPT3D", "ChannelPTExplicit", "ChannelPTSolver"]

PHI_BORDER = 4
PSI_BORDER = 2


def _csr(rows, cols, vals, n):
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sum_duplicates()
    A.eliminate_zeros()
    return A


def _block(B, row0, col0):
    B = sp.coo_matrix(B)
    return B.row + row0, B.col + col0, B.data.astype(float)


@dataclass
class ChannelPT3D:
    Nx: int
    Ny: int
    Nz: int
    Re: float
    Lx: float = 4 * np.pi
    Lz: float = 2 * np.pi
    G: float | None = None
    dealias: float = 1.5
    phi_mats: list = field(default_factory=list, repr=False)
    psi_mats: list = field(default_factory=list, repr=False)

    kind = "channel_pt"

    def __post_init__(self):
        if self.Nx % 2 or self.Nz % 2:
            raise ValueError("Nx and Nz must be even (the Nyquist modes are not represented)")
        if self.Ny < 6:
            raise ValueError("Ny must be >= 6 (four phi taus and two boundary conditions)")
        if self.G is None:
            self.G = 2.0 / self.Re
        self.Kx, self.Kzh, self.Kz = self.Nx // 2, self.Nz // 2, self.Nz - 1
        self.P = self.Kx * self.Kzh
        self.K = self.Kx * self.Kz                       # represented complex modes
        self.N = self.Ny
        self.nP, self.nS = self.N + 4, self.N + 2
        self.kx0, self.kz0 = 2 * np.pi / self.Lx, 2 * np.pi / self.Lz
        self.Mx = S.dealias_size(self.Nx, self.dealias)
        self.My = S.dealias_size(self.Ny, self.dealias)
        self.Mz = S.dealias_size(self.Nz, self.dealias)
        self.size_phi = 4 * self.P * self.nP
        self.size_psi = 4 * self.P * self.nS
        self.size = self.size_phi + self.size_psi
        N = self.N
        self.S04 = S.conversion_chain(N, 0, 4)
        self.S24 = S.conversion_chain(N, 2, 4)
        self.S02 = S.conversion_chain(N, 0, 2)
        self.S12 = S.conversion_chain(N, 1, 2)
        self.S14 = S.conversion_chain(N, 1, 4)
        self.D01 = S.differentiation(N, 0, 1)
        self.D02 = S.differentiation(N, 0, 2)
        self.D04 = S.differentiation(N, 0, 4)
        self.S2D = (self.S24 @ self.D02).tocsr()           # d^2 in C^(4)
        self._bc()
        self.phi_mats = [self.phi_matrices(s) for s in range(self.P)]
        self.psi_mats = [self.psi_matrices(s) for s in range(self.P)]

    # -- modes and slots ---------------------------------------------------------------------------
    def slot(self, kx: int, kz: int) -> tuple[int, int]:
        """(slot, column pair) of mode (kx, kz): columns (0, 1) for kz >= 0, (2, 3) for kz < 0."""
        return kx * self.Kzh + abs(kz), (0 if kz >= 0 else 2)

    def slot_wavenumbers(self, s: int) -> tuple[int, int]:
        return int(s // self.Kzh), int(s % self.Kzh)

    def k2(self, s: int) -> float:
        kx, kz = self.slot_wavenumbers(s)
        ax, az = kx * self.kx0, kz * self.kz0
        return ax * ax + az * az

    def kz_of(self, iz):
        iz = np.asarray(iz)
        return np.where(iz < self.Kzh, iz, iz - self.Kz)

    # -- operators ---------------------------------------------------------------------------------
    def _bc(self):
        n = np.arange(self.N, dtype=float)
        self.bc_value = ((-1.0) ** n, np.ones(self.N))                     # T_n(-1), T_n(1)
        self.bc_deriv = (-((-1.0) ** n) * n * n, n * n)                     # T_n'(-1), T_n'(1)

    def phi_matrices(self, s: int):
        """(M, L) of the phi system of slot s: real CSR of size N + 4, border rows 0..3."""
        N, n = self.N, self.nP
        Mr, Mc, Mv, Lr, Lc, Lv = [], [], [], [], [], []
        for r, vals in enumerate((self.bc_value[0], self.bc_value[1], self.bc_deriv[0], self.bc_deriv[1])):
            nz = np.flatnonzero(vals)
            Lr.append(np.full(nz.size, r))
            Lc.append(nz)
            Lv.append(vals[nz])
        if s == 0:
            Lb = self.S04                                                    # <phi> = 0 pin
        else:
            k2 = self.k2(s)
            Mb = (-k2) * self.S2D + (k2 * k2) * self.S04
            for a, b, c in [_block(Mb, 4, 0)]:
                Mr.append(a), Mc.append(b), Mv.append(c)
            Lb = (k2 / self.Re) * (self.D04 - (2.0 * k2) * self.S2D + (k2 * k2) * self.S04)
        a, b, c = _block(Lb, 4, 0)
        Lr.append(a), Lc.append(b), Lv.append(c)
        for k in range(1, 5):                                               # tau_k in row N - k
            Lr.append(np.array([4 + N - k]))
            Lc.append(np.array([N + 4 - k]))
            Lv.append(np.array([1.0]))
        M = _csr(Mr, Mc, Mv, n) if Mr else sp.csr_matrix((n, n))
        return M, _csr(Lr, Lc, Lv, n)

    def psi_matrices(self, s: int):
        """(M, L) of the psi system of slot s (slot 0: the mean flow U, W), size N + 2."""
        N, n = self.N, self.nS
        Mr, Mc, Mv, Lr, Lc, Lv = [], [], [], [], [], []
        for r, vals in enumerate(self.bc_value):
            nz = np.flatnonzero(vals)
            Lr.append(np.full(nz.size, r))
            Lc.append(nz)
            Lv.append(vals[nz])
        if s == 0:
            Mb = self.S02
            Lb = (-1.0 / self.Re) * self.D02
        else:
            k2 = self.k2(s)
            Mb = (-k2) * self.S02
            Lb = (k2 / self.Re) * (self.D02 - k2 * self.S02)
        for blk, (R, Cc, V) in ((Mb, (Mr, Mc, Mv)), (Lb, (Lr, Lc, Lv))):
            a, b, c = _block(blk, 2, 0)
            R.append(a), Cc.append(b), V.append(c)
        for k in range(1, 3):
            Lr.append(np.array([2 + N - k]))
            Lc.append(np.array([N + 2 - k]))
            Lv.append(np.array([1.0]))
        return _csr(Mr, Mc, Mv, n), _csr(Lr, Lc, Lv, n)

    @property
    def buckets(self):
        """(name, slots, n, border rows, offset into the state, matrices) of the two buckets."""
        return (("phi", self.P, self.nP, list(range(PHI_BORDER)), 0, self.phi_mats),
                ("psi", self.P, self.nS, list(range(PSI_BORDER)), self.size_phi, self.psi_mats))

    # -- state layout ------------------------------------------------------------------------------
    def split(self, X: np.ndarray):
        """Views (4, P, N + 4) and (4, P, N + 2) of the phi and psi buckets of a flat state."""
        X = np.asarray(X)
        return X[:self.size_phi].reshape(4, self.P, self.nP), X[self.size_phi:].reshape(4, self.P, self.nS)

    def to_modes(self, X: np.ndarray):
        """Complex (Kx, Kz, N + 4) phi and (Kx, Kz, N + 2) psi per mode (kz in FFT order), and the
        real mean-flow vectors U, W (N + 2, taus included); the (0, 0) phi and psi are zero."""
        A, B = self.split(X)
        phi = np.zeros((self.Kx, self.Kz, self.nP), complex)
        psi = np.zeros((self.Kx, self.Kz, self.nS), complex)
        kz = self.kz_of(np.arange(self.Kz))
        for iz, k in enumerate(kz):
            s = np.arange(self.Kx) * self.Kzh + abs(int(k))
            c = 0 if k >= 0 else 2
            phi[:, iz] = A[c, s] + 1j * A[c + 1, s]
            psi[:, iz] = B[c, s] + 1j * B[c + 1, s]
        U, W = B[0, 0].copy(), B[1, 0].copy()
        phi[0, 0] = 0.0
        psi[0, 0] = 0.0
        return phi, psi, U, W

    def from_modes(self, phi, psi, U, W) -> np.ndarray:
        """The flat state of per-mode phi, psi (kz in FFT order) and the mean flow U, W."""
        X = np.zeros(self.size)
        A, B = self.split(X)
        kz = self.kz_of(np.arange(self.Kz))
        for iz, k in enumerate(kz):
            s = np.arange(self.Kx) * self.Kzh + abs(int(k))
            c = 0 if k >= 0 else 2
            A[c, s], A[c + 1, s] = phi[:, iz].real, phi[:, iz].imag
            B[c, s], B[c + 1, s] = psi[:, iz].real, psi[:, iz].imag
        A[:, 0] = 0.0
        B[:, 0] = 0.0
        B[0, 0], B[1, 0] = np.asarray(U, float), np.asarray(W, float)
        return X

    def velocity(self, X: np.ndarray) -> np.ndarray:
        """(3, Kx, Kz, N) complex Chebyshev coefficients of u, v, w per mode."""
        phi, psi, U, W = self.to_modes(X)
        N = self.N
        ax = (self.kx0 * np.arange(self.Kx))[:, None, None]
        az = (self.kz0 * self.kz_of(np.arange(self.Kz)))[None, :, None]
        f, g = phi[..., :N], psi[..., :N]
        df = _cheb_diff_last(f)
        V = np.stack([1j * ax * df + 1j * az * g, (ax * ax + az * az) * f, 1j * az * df - 1j * ax * g])
        V[0, 0, 0] = U[:N]
        V[2, 0, 0] = W[:N]
        V[1, 0, 0] = 0.0
        return V

    def unknown_name(self, j: int) -> str:
        """Name of entry j of the flat state (bucket, column, slot, row)."""
        if j < self.size_phi:
            c, rest = divmod(j, self.P * self.nP)
            s, r = divmod(rest, self.nP)
            what = f"phi[{r}]" if r < self.N else f"tau_phi{self.N + 4 - r}"
        else:
            c, rest = divmod(j - self.size_phi, self.P * self.nS)
            s, r = divmod(rest, self.nS)
            what = (("U" if c == 0 else "W") if s == 0 else "psi") + (f"[{r}]" if r < self.N else f" tau{self.N + 2 - r}")
        kx, kz = self.slot_wavenumbers(s)
        return f"{what} of mode ({kx}, {'-' if c >= 2 else ''}{kz}), {'im' if c % 2 else 're'}"

    # -- states ------------------------------------------------------------------------------------
    def poiseuille(self) -> np.ndarray:
        """Laminar state U = (G Re / 2)(1 - y^2), everything else zero."""
        U = np.zeros(self.nS)
        a = self.G * self.Re / 2
        U[0], U[2] = a / 2, -a / 2                     # 1 - y^2 = T_0/2 - T_2/2
        phi = np.zeros((self.Kx, self.Kz, self.nP), complex)
        psi = np.zeros((self.Kx, self.Kz, self.nS), complex)
        return self.from_modes(phi, psi, U, np.zeros(self.nS))

    def initial_state(self, seed: int = 0, energy: float = 1e-4, kmax: int = 4) -> np.ndarray:
        """Poiseuille flow plus a seeded perturbation: phi = (1 - y^2)^2 a and psi = (1 - y^2) b
        with random complex a, b on the modes |kx|, |kz| <= kmax ((0, 0) excluded; conjugate
        pairs at kx = 0), scaled to the volume-averaged kinetic energy ``energy``.  The
        perturbation is divergence-free by construction and vanishes at the walls (no slip)."""
        rng = np.random.default_rng(seed)
        N = self.N
        pphi = np.zeros(N)
        pphi[:5] = np.array([3.0, 0.0, -4.0, 0.0, 1.0]) / 8.0          # (1 - y^2)^2
        ppsi = np.zeros(N)
        ppsi[0], ppsi[2] = 0.5, -0.5                                     # 1 - y^2
        phi = np.zeros((self.Kx, self.Kz, self.nP), complex)
        psi = np.zeros((self.Kx, self.Kz, self.nS), complex)
        for kx in range(min(kmax, self.Kx - 1) + 1):
            for kz in range(-min(kmax, self.Kzh - 1), min(kmax, self.Kzh - 1) + 1):
                if (kx, kz) == (0, 0) or (kx == 0 and kz < 0):
                    continue
                a, b = rng.standard_normal(2) + 1j * rng.standard_normal(2)
                iz = kz % self.Kz
                phi[kx, iz, :N] = a * pphi
                psi[kx, iz, :N] = b * ppsi
                if kx == 0:
                    phi[0, (-kz) % self.Kz, :N] = np.conj(a) * pphi
                    psi[0, (-kz) % self.Kz, :N] = np.conj(b) * ppsi
        zero = np.zeros(self.nS)
        pert = self.from_modes(phi, psi, zero, zero)
        e = self.kinetic_energy(self.velocity(pert))
        if e > 0:
            pert *= np.sqrt(energy / e)
        return self.poiseuille() + pert

    def kinetic_energy(self, V: np.ndarray) -> float:
        """1/2 volume average of |u|^2 from (3, Kx, Kz, N) coefficients (Parseval in x, z,
        Gauss-Chebyshev quadrature of the exact polynomial product in y)."""
        My = 2 * self.N
        B = S.cheb_backward_matrix(self.N, My)
        vals = np.einsum("jn,fxzn->fxzj", B, V)
        w = np.full(self.Kx, 2.0)
        w[0] = 1.0
        t = S.cheb_nodes(My)
        q = np.pi / My * np.sqrt(1 - t * t)
        return 0.5 * float(np.einsum("x,fxzj,j->", w, np.abs(vals) ** 2, q)) / 2.0

    # -- what the stepping machinery needs ---------------------------------------------------------
    def discretization_key(self) -> tuple:
        return ("channel_pt", self.Nx, self.Ny, self.Nz, float(self.Lx), float(self.Lz), float(self.dealias),
                float(self.G))

    def params(self) -> tuple:
        return (float(self.Re),)

    def describe(self) -> str:
        return f"channel 3D (poloidal-toroidal) {self.Nx}x{self.Ny}x{self.Nz}, Re={self.Re:g}"

    def identity(self) -> dict:
        return {"problem": "channel_3d_pt", "Nx": int(self.Nx), "Ny": int(self.Ny), "Nz": int(self.Nz),
                "Lx": float(self.Lx), "Lz": float(self.Lz), "dealias": float(self.dealias), "Re": float(self.Re),
                "G": float(self.G)}


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


# ------------------------------------------------------------------------------------------------
# device stepping
# ------------------------------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _diagonals(A, offsets, N):
    """(N, len(offsets)) table of A[m, m + o] (zero outside the matrix)."""
    A = A.tocsr()
    out = np.zeros((N, len(offsets)))
    for t, o in enumerate(offsets):
        m = np.arange(max(0, N - o))
        out[m, t] = np.asarray(A[m, m + o]).ravel()
    return out


def _cdiv(a: int, b: int) -> int:
    return -(-a // b)


def y_fold_geometry(N: int, My: int) -> dict:
    """Sizes of the parity-folded y tables (the same formulas as the kernels' ``Fold``): Mh folded
    nodes (MHP padded to 16, Mhs to 8), KS coefficient pairs (padded to 8), RE / RO rows of the
    even / odd row group (padded to 16); the tables carry one zero step more, KT = KS + 4 and
    MhT = Mhs + 4 rows, for the kernels' load-ahead."""
    Mh = (My + 1) // 2
    KS = 8 * _cdiv((N + 1) // 2, 8)
    Mhs = 8 * _cdiv(Mh, 8)
    return dict(Mh=Mh, MHP=16 * _cdiv(Mh, 16), Mhs=Mhs, MhT=Mhs + 4, KS=KS, KT=KS + 4,
                RE=16 * _cdiv((N + 1) // 2, 16), RO=16 * _cdiv(N // 2, 16))


def y_fold_tables(pb: "ChannelPT3D"):
    """The y tables of cn_chpt_explicit.  The Gauss nodes are symmetric (y_{My-1-j} = -y_j), so
    T_n and T'_n have parity (-1)^n and (-1)^(n+1) over the node pairs, and every row of the five
    forward matrices has one parity: (-1)^m for S04 Fy, S24 D02 Fy, S02 Fy and (-1)^(m+1) for
    S14 D01 Fy, S12 D01 Fy.  Both are checked here to 1e-12 of the largest entry before the
    first Mh node columns are kept.

    Returns (Tb[4][KT][MHP], Am[5][MhT][RE + RO], s02c0): Tb holds T_{2k}, T_{2k+1}, T'_{2k},
    T'_{2k+1} at the first Mh nodes; Am holds the forward matrices node-major with the rows
    permuted even first (row m = 2 i at i, m = 2 i + 1 at RE + i); zero padding throughout."""
    N, My = pb.N, pb.My
    g = y_fold_geometry(N, My)
    Mh = g["Mh"]
    T = S.cheb_backward_matrix(N, My, 0)
    D = S.cheb_backward_matrix(N, My, 1, 2.0)
    half = My // 2
    n = np.arange(N)
    for B, sg in ((T, (-1.0) ** n), (D, -(-1.0) ** n)):
        err = np.max(abs(B[::-1][:half] - sg * B[:half]))
        if not err <= 1e-12 * np.max(abs(B)):
            raise AssertionError(f"Chebyshev tables not parity-symmetric ({err:.2e})")
    Tb = np.zeros((4, g["KT"], g["MHP"]))
    for k in range(g["KS"]):
        for p, (B, nn) in enumerate(((T, 2 * k), (T, 2 * k + 1), (D, 2 * k), (D, 2 * k + 1))):
            if nn < N:
                Tb[p, k, :Mh] = B[:Mh, nn]
    Fy = S.cheb_forward_matrix(N, My)
    mats = [(pb.S14 @ pb.D01).toarray() @ Fy, pb.S2D.toarray() @ Fy, pb.S04.toarray() @ Fy,
            pb.S02.toarray() @ Fy, (pb.S12 @ pb.D01).toarray() @ Fy]
    odd = (0, 4)                                     # the D01 forms: parity (-1)^(m+1)
    m = np.arange(N)
    RP = g["RE"] + g["RO"]
    perm = np.where(m % 2 == 0, m // 2, g["RE"] + m // 2)
    Am = np.zeros((5, g["MhT"], RP))
    for t, A in enumerate(mats):
        sg = (-1.0) ** m * (-1.0 if t in odd else 1.0)
        err = np.max(abs(A[:, ::-1][:, :half] - sg[:, None] * A[:, :half]))
        if not err <= 1e-12 * np.max(abs(A)):
            raise AssertionError(f"forward matrix {t} not parity-symmetric ({err:.2e})")
        Am[t, :Mh, perm] = A[:, :Mh]
    return Tb, Am, pb.S02.toarray()[:, 0].copy()


class ChannelPTExplicit:
    """F(X) of the poloidal-toroidal channel on the device (cn_chpt_explicit: five launches, one
    workspace); the tables are built once per discretization."""

    def __init__(self, pb: ChannelPT3D):
        from . import _lib
        torch = _torch()
        self.pb = pb
        dev = "cuda"
        Tb, Am, s02c0 = y_fold_tables(pb)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)           # noqa: E731
        self.Tb, self.Am, self.s02c0 = t(Tb), t(Am), t(s02c0)
        self.Wx = t(np.exp(-2j * np.pi * np.arange(pb.Mx) / pb.Mx))
        self.Wz = t(np.exp(-2j * np.pi * np.arange(pb.Mz) / pb.Mz))
        nbytes = int(_lib.lib().cn_chpt_workspace_bytes(pb.Nx, pb.Ny, pb.Nz, pb.Mx, pb.My, pb.Mz))
        self.ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)

    def write(self, x, out) -> None:
        from . import _lib
        pb = self.pb
        v = ctypes.c_void_p
        p = lambda a: v(a.data_ptr())                                             # noqa: E731
        st = _lib.lib().cn_chpt_explicit(p(x), p(out), p(self.ws), self.ws.numel(), p(self.Tb), p(self.Am),
                                         p(self.s02c0), p(self.Wx), p(self.Wz), pb.Nx, pb.Ny, pb.Nz, pb.Mx, pb.My,
                                         pb.Mz, float(pb.kx0), float(pb.kz0), float(pb.G),
                                         v(_torch().cuda.current_stream().cuda_stream))
        _lib.check(st, "ChannelPTExplicit")


class ChannelPTSolver:
    """IVP solver of the poloidal-toroidal channel on the device: SBDF1..4 with the lower-order
    startup (the reference's IVPSolver step, python/conatus/timestep/integrator.py:648-682,
    361-373).  One step is the explicit term (six launches) written into the newest history
    level, the fused stage (M X and the right-hand side over X in place) and one in-place solve
    of each bucket: the 4 P phi and 4 P psi right-hand sides against P real factorizations each
    (the modes (kx, kz) and (kx, -kz) share theirs).  ``run_graph`` replays full-order steps from
    captured HIP graphs: one per whole turn of the history ring (order steps) while at least a
    turn is left, one per ring phase for the rest.  The clock is t_origin + (iteration -
    it_origin) dt, and the state is checked finite every ``check_finite`` steps (``run``) or at
    the end of the call (``run_graph``).""" 

    def __init__(self, problem: ChannelPT3D, scheme: str, dt: float, pivot: bool = True, mode: str | None = None,
                 check_finite=True):
        from .errors import OptionError
        from .runtime.imex import sbdf_coefficients
        from .timestep import finite_cadence
        torch = _torch()
        if scheme not in ("SBDF1", "SBDF2", "SBDF3", "SBDF4"):
            raise OptionError(f"ChannelPTSolver: scheme {scheme!r} not supported", fix="one of SBDF1..SBDF4")
        if not float(dt) > 0:
            raise OptionError("dt must be > 0")
        self.problem = pb = problem
        self.scheme, self.dt, self.order = scheme, float(dt), int(scheme[4])
        self._coef = sbdf_coefficients
        self.pivot, self.mode = bool(pivot), mode
        self._finite_every = finite_cadence(check_finite, f"ChannelPTSolver(problem, {scheme!r}, ...)")
        self.explicit = ChannelPTExplicit(pb)
        N = pb.N
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")        # noqa: E731
        self._s04 = t(_diagonals(pb.S04, (0, 2, 4, 6, 8), N))
        self._s2d = t(_diagonals(pb.S2D, (0, 2, 4, 6, 8), N))
        self._s02 = t(_diagonals(pb.S02, (0, 2, 4), N))
        z = lambda: torch.zeros(pb.size, dtype=torch.float64, device="cuda")       # noqa: E731
        self.X = z()
        self._ring = [(z(), z()) for _ in range(self.order)]        # (M X, F) levels
        self._factors: dict = {}
        self._graphs: dict = {}
        self.t, self.t_origin, self.it_origin = 0.0, 0.0, 0
        self._finite_ok = None
        self.reset()

    # -- state -------------------------------------------------------------------------------------
    def reset(self) -> None:
        self._levels: list[int] = []
        self._last = None
        self.iteration = 0

    @property
    def history_length(self) -> int:
        return len(self._levels)

    @property
    def n_factorizations(self) -> int:
        return len(self._factors)

    def set_state(self, X: np.ndarray, t: float = 0.0) -> None:
        X = np.asarray(X, dtype=np.float64).ravel()
        if X.size != self.problem.size:
            from .errors import OptionError
            raise OptionError(f"ChannelPTSolver.set_state: {self.problem.size} values expected, got {X.size}")
        self.X.copy_(_torch().from_numpy(np.ascontiguousarray(X)))
        self.reset()
        self.t, self.t_origin, self.it_origin = float(t), float(t), 0
        self._finite_ok = None

    def state(self) -> np.ndarray:
        return self.X.cpu().numpy()

    def _clock(self) -> float:
        return self.t_origin + (self.iteration - self.it_origin) * self.dt

    # -- operators ---------------------------------------------------------------------------------
    def _factor(self, alpha: float):
        f = self._factors.get(alpha)
        if f is None:
            from .runtime.imex import factor_operator
            f = []
            for name, P, n, border, off, mats in self.problem.buckets:
                h = factor_operator([m[0] for m in mats], [m[1] for m in mats], alpha, 1.0, border, np.float64,
                                    self.pivot, self.mode)
                h.reserve(4)
                f.append((h, off, P * n))
            self._factors[alpha] = f
        return f

    # -- steps -------------------------------------------------------------------------------------
    def _next_levels(self, levels=None):
        levels = self._levels if levels is None else levels
        if len(levels) < self.order:
            used = set(levels)
            slot = next(i for i in range(self.order) if i not in used)
            return [slot] + levels
        return [levels[-1]] + levels[:-1]

    def _launch(self, levels) -> None:
        from . import _lib
        torch = _torch()
        pb = self.problem
        q = len(levels)
        a, c = self._coef(q)
        mx0, f0 = self._ring[levels[0]]
        self.explicit.write(self.X, f0)
        P = ctypes.c_void_p * 4
        fl = P(*([self._ring[i][1].data_ptr() for i in levels] + [0] * (4 - q)))
        mh = P(*([self._ring[i][0].data_ptr() for i in levels[1:]] + [0] * (5 - q)))
        cc = (ctypes.c_double * 5)(*([0.0] + [float(c[j]) for j in range(1, q + 1)] + [0.0] * (4 - q)))
        adt = (ctypes.c_double * 5)(*([0.0] + [float(a[j] / self.dt) for j in range(1, q + 1)] + [0.0] * (4 - q)))
        v = ctypes.c_void_p
        stream = torch.cuda.current_stream().cuda_stream
        st = _lib.lib().cn_chpt_sbdf_stage(v(self.X.data_ptr()), fl, v(mx0.data_ptr()), mh, q, cc, adt,
                                           v(self._s04.data_ptr()), v(self._s2d.data_ptr()), v(self._s02.data_ptr()),
                                           pb.Nx, pb.Ny, pb.Nz, float(pb.kx0), float(pb.kz0), v(stream))
        _lib.check(st, "ChannelPTSolver (stage)")
        # the buckets one after the other on this stream: each solve fills the CUs with its own
        # LDS-sized workgroups, so two streams would only add fork and join gaps
        for h, off, st in self._factor(a[0] / self.dt):
            h.solve_ptr(self.X.data_ptr() + 8 * off, 4, st, False, stream)

    def _continuation(self) -> None:
        torch = _torch()
        if self._levels and not (self._last is not None and self._last[1] == self.t
                                 and torch.equal(self._last[0].view(torch.uint8), self.X.view(torch.uint8))):
            from .errors import OptionError
            raise OptionError("ChannelPTSolver: the history belongs to another state (the state was changed "
                              "outside the solver)", fix="set_state(X, t) starts a new trajectory (reset)")

    def _remember(self) -> None:
        self._last = (self.X.clone(), float(self.t)) if self._levels else None

    def _eager(self, steps: int) -> None:
        for _ in range(steps):
            levels = self._next_levels()
            self._launch(levels)
            self._levels = levels
            self.iteration += 1
            self.t = self._clock()

    def step(self) -> None:
        self.run(1)

    def run(self, steps: int) -> float:
        self._continuation()
        fe, left = self._finite_every, int(steps)
        while left:
            k = min(left, fe - self.iteration % fe) if fe else left
            self._eager(k)
            left -= k
            if fe and self.iteration % fe == 0 and left:
                self.check_finite("ChannelPTSolver.run")
        if fe and steps:
            self.check_finite("ChannelPTSolver.run (end of the call)")
        self._remember()
        return self.t

    def run_graph(self, steps: int) -> float:
        from .runtime.imex import capture_graph
        self._continuation()
        done = 0
        while self.history_length < self.order - 1 and done < steps:
            self._eager(1)
            done += 1
        if done < steps:
            a, _ = self._coef(self.order)
            self._factor(a[0] / self.dt)
        while done < steps:
            if len(self._levels) == self.order and steps - done >= self.order:
                # a whole turn of the history ring (order steps) in one graph: one replay
                # instead of order, so the launch gaps between replays fall away
                seq, lv = [], self._levels
                for _ in range(self.order):
                    lv = self._next_levels(lv)
                    seq.append(lv)
                key = ("turn",) + tuple(seq[0])
                g = self._graphs.get(key)
                if g is None:
                    g = capture_graph(lambda sq=seq: [self._launch(x) for x in sq])
                    self._graphs[key] = g
                g.replay()
                self._levels = seq[-1]
                self.iteration += self.order
                self.t = self._clock()
                done += self.order
                continue
            levels = self._next_levels()
            key = tuple(levels)
            g = self._graphs.get(key)
            if g is None:
                g = capture_graph(lambda lv=levels: self._launch(lv))
                self._graphs[key] = g
            g.replay()
            self._levels = levels
            self.iteration += 1
            self.t = self._clock()
            done += 1
        if self._finite_every and steps:
            self.check_finite("ChannelPTSolver.run_graph (end of the call)")
        self._remember()
        return self.t

    def check_finite(self, where: str = "ChannelPTSolver.check_finite") -> None:
        torch = _torch()
        if bool(torch.isfinite(self.X).all()):
            self._finite_ok = (self.iteration, self.t)
            return
        from .errors import NumericalError
        bad = ~torch.isfinite(self.X).cpu().numpy()
        j = int(np.flatnonzero(bad)[0])
        last = self._finite_ok
        since = f"; last checked finite at iteration {last[0]} (t = {last[1]:.10g})" if last else ""
        raise NumericalError(f"non-finite state (NaN or inf) at iteration {self.iteration}, t = {self.t:.10g}: "
                             f"{self.problem.unknown_name(j)} ({int(bad.sum())} non-finite values){since}",
                             where=where, fix=f"reduce dt (now {self.dt:g}) or raise the resolution")

    def explain(self) -> str:
        pb = self.problem
        return (f"ChannelPTSolver({pb.describe()}, {self.scheme}, dt={self.dt:g}): {pb.K} complex modes as "
                f"{pb.P} real phi systems (n={pb.nP}, kl=ku=4, r=4) and {pb.P} real psi systems (n={pb.nS}, "
                f"kl=ku=2, r=2), four right-hand sides each; F from six products on the {pb.Mx}x{pb.My}x{pb.Mz} grid")
