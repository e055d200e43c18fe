"""CPU oracle of the 2D Rayleigh-Benard IMEX step — TEST INFRASTRUCTURE ONLY (see oracle/banded.py).

Consumes the per-mode operators (M_k, L_k), border rows and constant rows b0 built by
conatus_amd.problems.RayleighBenard2D (the discretization is data here; it is pinned separately
by the linear growth rate 5.5425059 at Ra = 2500, k = pi, Pr = 1 and by Ra_c = 1707.762, the
reference's validation values, python/conatus/problems/convection.py:38-40).  Everything else is
restated from the reference, independently of the device kernels:

* explicit term as the reference's evaluator path (python/conatus/timestep/spectral.py:53-82,
  223-250 and python/conatus/reference/transforms.py:103-177): complex modes converted to the
  RealFourier pair storage [cos_k, msin_k], x-derivative by the pair rule out_cos = -k msin,
  out_msin = k cos, z-derivative by the Chebyshev backward recurrence, FFT-based backward
  transforms to the 3/2 grid, the products -u.grad(u), -u.grad(T) on the grid, forward
  transforms (truncation = dealiasing projection), conversion S_{0->2} into the equation rows;
* SBDF step of core/include/conatus/imex.hpp:158-176 per mode, and the IMEX Runge-Kutta stages
  of python/conatus/timestep/integrator.py:608-646 (stage j's F and L X only if a later stage
  uses them; rhs = M X0 + sum dt A F - dt H L X in that order), solves by oracle.banded (LAPACK).
"""
from __future__ import annotations

import numpy as np

from . import spectral as S
from .banded import BorderedBanded
from .imex import CN_SCHEMES, cn_coefficients, cn_rhs, rk_tableau, sbdf_coefficients


def to_pairs(c: np.ndarray) -> np.ndarray:
    """(K, N) complex modes -> (2K, N) RealFourier pairs [cos_0, msin_0, cos_1, msin_1, ...]."""
    K, N = c.shape
    p = np.zeros((2 * K, N))
    p[0] = c[0].real
    p[2::2] = 2 * c[1:].real
    p[3::2] = 2 * c[1:].imag
    return p


def from_pairs(p: np.ndarray) -> np.ndarray:
    K = p.shape[0] // 2
    c = np.zeros((K, p.shape[1]), complex)
    c[0] = p[0]
    c[1:] = (p[2::2] + 1j * p[3::2]) / 2
    return c


def dx_pairs(p: np.ndarray, kappa: float) -> np.ndarray:
    K = p.shape[0] // 2
    k = (np.arange(K) * kappa)[:, None]
    out = np.zeros_like(p)
    out[0::2] = -k * p[1::2]
    out[1::2] = k * p[0::2]
    return out


class RB2DOracle:
    def __init__(self, pb, scheme: str = "SBDF2", dt: float = 1e-3):
        self.pb = pb
        self.order = int(scheme[4]) if scheme.startswith("SBDF") else 0
        self.cn = cn_coefficients(scheme) if scheme in CN_SCHEMES else None
        self.tab = None if (self.order or self.cn is not None) else rk_tableau(scheme)
        self.LXh = []
        self.dt = dt
        self.M = [m[0] for m in pb.mats]
        self.L = [m[1] for m in pb.mats]
        self.b0 = pb.b0()
        self.MXh, self.Fh = [], []
        self.factors = {}

    def _grid(self, pairs):
        pb = self.pb
        g = S.rf_backward(pairs, pb.Mx, axis=0)
        return S.cheb_backward(g, pb.Mz, axis=1)

    def explicit(self, X: np.ndarray) -> np.ndarray:
        pb = self.pb
        fld = lambda f: X[:, pb.field_index(f)]                         # noqa: E731
        U, W, T = (to_pairs(fld(f)) for f in (0, 1, 2))
        u, w = self._grid(U), self._grid(W)
        ux, wx, Tx = (self._grid(dx_pairs(A, pb.kappa)) for A in (U, W, T))
        uz, wz, Tz = (self._grid(S.cheb_diff(A, pb.Lz, axis=1)) for A in (U, W, T))
        outs = (-(u * ux + w * uz), -(u * wx + w * wz), -(u * Tx + w * Tz))
        F = self.b0.copy()
        for e, g in enumerate(outs):
            c = S.cheb_forward(S.rf_forward(g, pb.Nx, axis=0), pb.N, axis=1)    # (Nx pairs, N)
            rows = (pb.S02 @ from_pairs(c).T).T                                  # (K, N) in C^(2)
            F[:, pb.equation_rows(e)] = rows
        return F

    def factor(self, k, alpha, beta=1.0):
        key = (k, alpha, beta)
        if key not in self.factors:
            A = (np.complex128(alpha) * self.M[k] + np.complex128(beta) * self.L[k]).tocsr()
            self.factors[key] = BorderedBanded(A, self.pb.border, np.complex128)
        return self.factors[key]

    def step(self, X: np.ndarray) -> np.ndarray:
        K = self.pb.K
        MX = np.stack([self.M[k] @ X[k] for k in range(K)])
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
        return np.stack([self.factor(k, a[0] / self.dt).solve(rhs[k]) for k in range(K)])

    def step_rk(self, X: np.ndarray) -> np.ndarray:
        K, dt = self.pb.K, self.dt
        A, H, cc = self.tab
        s = len(cc) - 1
        Xs, F, LX = [X] + [None] * s, [None] * (s + 1), [None] * (s + 1)
        MX0 = np.stack([self.M[k] @ X[k] for k in range(K)])
        for i in range(1, s + 1):
            j = i - 1
            if any(A[m][j] != 0 for m in range(j + 1, s + 1)):
                F[j] = self.explicit(Xs[j])
            if any(H[m][j] != 0 for m in range(j + 1, s + 1)):
                LX[j] = np.stack([self.L[k] @ Xs[j][k] for k in range(K)])
            rhs = MX0.copy()
            for k in range(i):
                if A[i][k] != 0:
                    rhs = rhs + (dt * A[i][k]) * F[k]
                if H[i][k] != 0:
                    rhs = rhs - (dt * H[i][k]) * LX[k]
            Xs[i] = np.stack([self.factor(kk, 1.0, dt * H[i][i]).solve(rhs[kk]) for kk in range(K)])
        return Xs[s]

    def step_cn(self, X: np.ndarray) -> np.ndarray:
        """CN-family multistep step per mode (oracle.imex.ImexIntegrator.step_cn)."""
        K = self.pb.K
        s = len(self.cn[0]) - 1
        self.MXh.insert(0, np.stack([self.M[k] @ X[k] for k in range(K)]))
        self.LXh.insert(0, np.stack([self.L[k] @ X[k] for k in range(K)]))
        self.Fh.insert(0, self.explicit(X))
        while len(self.MXh) > s:
            self.MXh.pop()
            self.LXh.pop()
            self.Fh.pop()
        a, b, c = self.cn if len(self.MXh) >= s else cn_coefficients("CNAB1")
        rhs = cn_rhs(a, b, c, self.dt, self.MXh, self.LXh, self.Fh)
        return np.stack([self.factor(k, a[0] / self.dt, b[0]).solve(rhs[k]) for k in range(K)])

    def run(self, X: np.ndarray, steps: int) -> np.ndarray:
        for _ in range(steps):
            if self.order:
                X = self.step(X)
            elif self.cn is not None:
                X = self.step_cn(X)
            else:
                X = self.step_rk(X)
        return X
