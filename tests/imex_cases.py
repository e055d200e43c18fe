"""IMEX test problems restated from the reference's tests: the manufactured forced heat equation
on a Chebyshev tau system (tests/core/test_core_imex.py:17-53) and the batched heat-like systems
of the cross-backend test (tests/cross_backend/test_gpu_vs_cpu.py:133-150)."""
import numpy as np
import scipy.sparse as sp

from oracle import spectral as S

N = 24
XG = S.cheb_grid(N)
S02 = S.conversion_chain(N, 0, 2)
D2 = S.differentiation(N, 0, 2)


def heat_system(N=N, kappa=0.0):
    """u_t - u_xx (+ kappa u) + lift taus = F, u(+-1) = 0; state [u_T (N), tau1, tau2]."""
    n = N + 2
    Sm = S.conversion_chain(N, 0, 2)
    M = sp.lil_matrix((n, n))
    L = sp.lil_matrix((n, n))
    M[:N, :N] = Sm
    L[:N, :N] = -S.differentiation(N, 0, 2) + kappa * Sm
    L[N - 1, N] = 1.0
    L[N - 2, N + 1] = 1.0
    L[N, :N] = (-1.0) ** np.arange(N)
    L[N + 1, :N] = 1.0
    return M.tocsr(), L.tocsr(), [N, N + 1]


PHI = {"cos": (np.cos, lambda t: -np.sin(t)),
       "flat": (lambda t: 1.0 + t ** 5, lambda t: 5.0 * t ** 4)}


def exact(t, phi="cos"):
    return np.sin(np.pi * XG) * PHI[phi][0](t)


def make_F(phi="cos"):
    f, df = PHI[phi]

    def F(x, t):
        u = S.cheb_backward(x[:N], N)
        g = np.sin(np.pi * XG) * (df(t) + (np.pi ** 2 + 1) * f(t))
        out = np.zeros(N + 2)
        out[:N] = S02 @ S.cheb_forward(-u + g, N)
        return out
    return F
