"""CPU oracle of the IMEX integrator — TEST INFRASTRUCTURE ONLY (see oracle/banded.py).

Restates core/include/conatus/imex.hpp (the reference's compiled ``ImexIntegrator``) in numpy,
operation by operation, with every product and sum rounded separately:

* scheme constants: ``sbdf_coefficients`` (imex.hpp:63-71), ``rk_tableau`` (imex.hpp:31-51),
  ``sbdf_order`` validation (imex.hpp:55-60);
* SBDF step (imex.hpp:158-176): push M X and F(X, t) to the history, q = min(order, levels),
  rhs = sum_{j=1..q} (c_j F_{j-1} - (a_j/dt) MX_{j-1}) accumulated from zero, solve
  (a_0/dt M + L) X = rhs;
* RK step (imex.hpp:178-205): stage i solves (M + dt H_ii L) X_i = M X_0 + dt sum_{k<i}
  (A_ik F_k - H_ik L X_k), F and L X evaluated only when a later stage needs them; the tableaux
  RKSMR and RKGFY of python/conatus/timestep/schemes.py:122-137 in the same cumulative form;
* CN-family multistep step (python/conatus/timestep/integrator.py:576-602, coefficients
  schemes.py:91-103): push M X, L X and F(X, t); CNAB1 until the history is full; rhs =
  sum_{j>=1} (c_j F_j - (a_j/dt) M X_j - b_j L X_j) accumulated from zero in that order (zero
  coefficients skipped), solve (a_0/dt M + b_0 L) X = rhs;
* factor cache keyed by (alpha, beta) with LHS = alpha M + beta L (imex.hpp:143-152);
* continuation rule (imex.hpp:105-111): a second run must start from the returned (x, t).

The CSR products are scipy's ``A @ x`` (row sums accumulated from zero in storage order, as
``Csr::matvec``, banded.hpp:70-76).  Pinned by the closed-form tests of tests/test_oracle_imex.py
(observed temporal orders on a manufactured heat equation, as tests/core/test_core_imex.py:57-87,
the exact scheme constants, factor counts).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .banded import BorderedBanded


def sbdf_order(scheme: str) -> int:
    tail = scheme[4:]
    if not scheme.startswith("SBDF") or len(tail) != 1 or tail not in "1234":
        raise ValueError(f"unknown SBDF scheme '{scheme}': SBDF1, SBDF2, SBDF3 or SBDF4")
    return int(tail)


def sbdf_coefficients(order: int):
    """(a, c) with a scaled by dt (imex.hpp:63-71)."""
    table = {1: ([1.0, -1.0], [0.0, 1.0]),
             2: ([1.5, -2.0, 0.5], [0.0, 2.0, -1.0]),
             3: ([11.0 / 6, -3.0, 1.5, -1.0 / 3], [0.0, 3.0, -3.0, 1.0]),
             4: ([25.0 / 12, -4.0, 3.0, -4.0 / 3, 0.25], [0.0, 4.0, -6.0, 4.0, -1.0])}
    if order not in table:
        raise ValueError("SBDF order must be 1..4")
    return table[order]


CN_SCHEMES = ("CNAB1", "CNAB2", "MCNAB2", "CNLF2")


def cn_coefficients(name: str):
    """(a, b, c), a in units of 1/dt (schemes.py:91-103, Wang & Ruuth 2008 eqs. 2.5, 2.9-2.11)."""
    table = {"CNAB1": ([1.0, -1.0], [0.5, 0.5], [0.0, 1.0]),
             "CNAB2": ([1.0, -1.0, 0.0], [0.5, 0.5, 0.0], [0.0, 1.5, -0.5]),
             "MCNAB2": ([1.0, -1.0, 0.0], [9 / 16, 6 / 16, 1 / 16], [0.0, 1.5, -0.5]),
             "CNLF2": ([0.5, 0.0, -0.5], [0.5, 0.0, 0.5], [0.0, 1.0, 0.0])}
    if name not in table:
        raise ValueError(f"unknown CN scheme: {name}")
    return table[name]


def rk_tableau(name: str):
    """(A, H, c) of the ARS tableaux (imex.hpp:31-51) and RKSMR, RKGFY (schemes.py:122-137)."""
    if name == "RK111":
        return [[0, 0], [1, 0]], [[0, 0], [0, 1]], [0, 1]
    if name == "RK222":
        g = (2.0 - np.sqrt(2.0)) / 2.0
        d = 1.0 - 1.0 / g / 2.0
        return ([[0, 0, 0], [g, 0, 0], [d, 1 - d, 0]], [[0, 0, 0], [0, g, 0], [0, 1 - g, g]], [0, g, 1])
    if name == "RK443":
        return ([[0, 0, 0, 0, 0], [1 / 2, 0, 0, 0, 0], [11 / 18, 1 / 18, 0, 0, 0],
                 [5 / 6, -5 / 6, 1 / 2, 0, 0], [1 / 4, 7 / 4, 3 / 4, -7 / 4, 0]],
                [[0, 0, 0, 0, 0], [0, 1 / 2, 0, 0, 0], [0, 1 / 6, 1 / 2, 0, 0],
                 [0, -1 / 2, 1 / 2, 1 / 2, 0], [0, 3 / 2, -3 / 2, 1 / 2, 1 / 2]],
                [0, 1 / 2, 2 / 3, 1 / 2, 1])
    if name == "RKSMR":
        # Spalart-Moser-Rogers substeps X_i = X_{i-1} + dt [g_i F_{i-1} + z_i F_{i-2} - a_i L X_{i-1}
        # - b_i L X_i], summed into cumulative rows
        al, be = (29 / 96, -3 / 40, 1 / 6), (37 / 160, 5 / 24, 1 / 6)
        ga, ze = (8 / 15, 5 / 12, 3 / 4), (-17 / 60, -5 / 12)
        A = [[0, 0, 0, 0], [ga[0], 0, 0, 0], [ga[0] + ze[0], ga[1], 0, 0],
             [ga[0] + ze[0], ga[1] + ze[1], ga[2], 0]]
        H = [[0, 0, 0, 0], [al[0], be[0], 0, 0], [al[0], be[0] + al[1], be[1], 0],
             [al[0], be[0] + al[1], be[1] + al[2], be[2]]]
        return A, H, [0, 8 / 15, 2 / 3, 1]
    if name == "RKGFY":                  # Heun on F, Crank-Nicolson on L
        return [[0, 0, 0], [1, 0, 0], [1 / 2, 1 / 2, 0]], [[0, 0, 0], [1 / 2, 1 / 2, 0], [1 / 2, 0, 1 / 2]], [0, 1, 1]
    raise ValueError(f"unknown RK scheme: {name}")


class ImexIntegrator:
    def __init__(self, M, L, border, scheme, dt, F=None, dtype=None):
        self.M, self.L = sp.csr_matrix(M), sp.csr_matrix(L)
        if self.M.shape != self.L.shape:
            raise ValueError("M and L sizes differ")
        if not dt > 0:
            raise ValueError("dt must be > 0")
        if dtype is None:
            dtype = np.complex128 if (np.iscomplexobj(self.M.data) or np.iscomplexobj(self.L.data)) else np.float64
        self.dtype = np.dtype(dtype)
        self.M = self.M.astype(self.dtype)
        self.L = self.L.astype(self.dtype)
        self.n = self.M.shape[0]
        self.border = list(border)
        self.scheme, self.dt, self.F = scheme, float(dt), F
        self.cn = cn_coefficients(scheme) if scheme in CN_SCHEMES else None
        if scheme.startswith("SBDF"):
            self.order = sbdf_order(scheme)
            self.tab = None
        elif self.cn is not None:
            self.order = 0
            self.tab = None
        else:
            self.order = 0
            self.tab = rk_tableau(scheme)
        self.factors = {}
        self.reset()

    def reset(self):
        self.MXh, self.Fh, self.LXh = [], [], []
        self.last = None
        self.iteration = 0

    @property
    def n_factorizations(self):
        return len(self.factors)

    @property
    def history_length(self):
        return len(self.MXh)

    def factor(self, alpha, beta):
        key = (alpha, beta)
        if key not in self.factors:
            lhs = (self.dtype.type(alpha) * self.M + self.dtype.type(beta) * self.L).tocsr()
            self.factors[key] = BorderedBanded(lhs, self.border, self.dtype)
        return self.factors[key]

    def evalF(self, x, t):
        if self.F is None:
            return np.zeros(self.n, self.dtype)
        return np.asarray(self.F(x, t), dtype=self.dtype)

    def run(self, x, t, nsteps):
        x = np.array(x, dtype=self.dtype)
        if self.MXh and not (self.last is not None and self.last[1] == t and np.array_equal(self.last[0], x)):
            raise ValueError("ImexIntegrator.run: history of a previous run; call reset()")
        for _ in range(nsteps):
            if self.order:
                x = self.step_sbdf(x, t)
            elif self.cn is not None:
                x = self.step_cn(x, t)
            else:
                x = self.step_rk(x, t)
            t += self.dt
            self.iteration += 1
        if self.order or self.cn is not None:
            self.last = (x.copy(), t)
        return x, t

    def step_sbdf(self, x, t):
        MX = self.M @ x
        FX = self.evalF(x, t)
        self.MXh.insert(0, MX)
        self.Fh.insert(0, FX)
        if len(self.MXh) > self.order:
            self.MXh.pop()
            self.Fh.pop()
        q = min(self.order, len(self.MXh))
        a, c = sbdf_coefficients(q)
        rhs = np.zeros(self.n, self.dtype)
        for j in range(1, q + 1):
            rhs = rhs + (c[j] * self.Fh[j - 1] - (a[j] / self.dt) * self.MXh[j - 1])
        return self.factor(a[0] / self.dt, 1.0).solve(rhs)

    def step_cn(self, x, t):
        s = len(self.cn[0]) - 1
        self.MXh.insert(0, self.M @ x)
        self.LXh.insert(0, self.L @ x)
        self.Fh.insert(0, self.evalF(x, t))
        while len(self.MXh) > s:
            self.MXh.pop()
            self.LXh.pop()
            self.Fh.pop()
        a, b, c = self.cn if len(self.MXh) >= s else cn_coefficients("CNAB1")
        return self.factor(a[0] / self.dt, b[0]).solve(cn_rhs(a, b, c, self.dt, self.MXh, self.LXh, self.Fh))

    def step_rk(self, x, t):
        A, H, cc = self.tab
        s = len(cc) - 1
        X = [None] * (s + 1)
        F = [None] * (s + 1)
        LX = [None] * (s + 1)
        X[0] = x.copy()
        MX0 = self.M @ x
        needF = lambda j: any(A[i][j] != 0 for i in range(j + 1, s + 1))    # noqa: E731
        needLX = lambda j: any(H[i][j] != 0 for i in range(j + 1, s + 1))   # noqa: E731
        for i in range(1, s + 1):
            j = i - 1
            if needF(j):
                F[j] = self.evalF(X[j], t + cc[j] * self.dt)
            if needLX(j):
                LX[j] = self.L @ X[j]
            rhs = MX0.copy()
            for k in range(i):
                if A[i][k] != 0:
                    rhs = rhs + (self.dt * A[i][k]) * F[k]
                if H[i][k] != 0:
                    rhs = rhs - (self.dt * H[i][k]) * LX[k]
            X[i] = self.factor(1.0, self.dt * H[i][i]).solve(rhs)
        return X[s]


def cn_rhs(a, b, c, dt, MXh, LXh, Fh):
    """sum_{j>=1} (c_j F_j - (a_j/dt) M X_j - b_j L X_j) from zero, in that order, zero
    coefficients skipped (integrator.py:581-594); MXh[j-1] = M X_j (newest first)."""
    rhs = np.zeros_like(MXh[0])
    for j in range(1, len(a)):
        if c[j] != 0:
            rhs = rhs + c[j] * Fh[j - 1]
        if a[j] != 0:
            rhs = rhs - (a[j] / dt) * MXh[j - 1]
        if b[j] != 0:
            rhs = rhs - b[j] * LXh[j - 1]
    return rhs
