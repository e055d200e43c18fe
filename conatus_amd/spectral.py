"""Spectral operators and transform matrices of the Fourier x Chebyshev IMEX path (host setup).

Conventions (python/conatus/reference/transforms.py:1-20,86-177; discretize/ultraspherical.py:43-90):
* Chebyshev axis on [a, b]: Gauss-Chebyshev nodes in ascending order t_j = -cos(pi (j + 1/2) / M),
  z = a + (t + 1) L / 2; coefficients of T_n, n < N; d/dz = (2/L) d/dt.
* Olver-Townsend ultraspherical operators, N x N, top modes dropped: conversion S_lam
  C^(lam) -> C^(lam+1), differentiation D_lam: C^(lam) -> C^(lam+1).
* Fourier axis of length Lx, M points x_j = a + Lx j / M, complex modes k = 0..K-1 (K = N/2, the
  Nyquist mode not represented): f(x) = u_0 + 2 Re sum_{k>=1} u_k e^{i k kappa (x - a)},
  kappa = 2 pi / Lx; u_k = (1/M) sum_j f_j e^{-i k kappa (x_j - a)} (numpy's rfft / M).  In the
  reference's RealFourier pair storage [cos_k, msin_k], u_0 = cos_0 and u_k = (cos_k + i msin_k)/2.
* Dealiased grids: M = floor(N * dealias) (dealias 3/2 makes quadratic products exact).

The transform matrices below are what the device kernels apply (a backward transform is one
matrix-vector product per column, derivatives folded into the matrix)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def conversion(N: int, lam: int) -> sp.csr_matrix:
    """S_lam: C^(lam) -> C^(lam+1); lam = 0 is T -> C^(1) (ultraspherical.py:43-56)."""
    n = np.arange(N, dtype=float)
    if lam == 0:
        d0 = np.full(N, 0.5)
        d0[0] = 1.0
        d2 = np.full(max(N - 2, 0), -0.5)
    else:
        d0 = lam / (n + lam)
        d2 = -lam / (n[2:] + lam)
    return sp.diags([d0, d2], [0, 2], shape=(N, N), format="csr")


def conversion_chain(N: int, lam_from: int, lam_to: int) -> sp.csr_matrix:
    S = sp.identity(N, format="csr")
    for lam in range(lam_from, lam_to):
        S = (conversion(N, lam) @ S).tocsr()
    return S


def differentiation(N: int, lam: int, order: int = 1) -> sp.csr_matrix:
    """d^order/dt^order: C^(lam) -> C^(lam+order) (ultraspherical.py:71-78)."""
    D = sp.identity(N, format="csr")
    for k in range(lam, lam + order):
        v = np.arange(1, N, dtype=float) if k == 0 else np.full(N - 1, 2.0 * k)
        D = (sp.diags([v], [1], shape=(N, N), format="csr") @ D).tocsr()
    return D


def dealias_size(N: int, dealias: float) -> int:
    return int(np.floor(N * dealias))


def cheb_nodes(M: int) -> np.ndarray:
    return -np.cos(np.pi * (np.arange(M) + 0.5) / M)


def cheb_grid(M: int, a: float, b: float) -> np.ndarray:
    return a + (cheb_nodes(M) + 1.0) * (b - a) / 2.0


def cheb_backward_matrix(N: int, M: int, deriv: int = 0, length: float = 2.0) -> np.ndarray:
    """(M, N): values (deriv = 0) or first derivatives (deriv = 1, d/dz on an interval of
    ``length``) of T_n at the M ascending Gauss-Chebyshev nodes."""
    theta = np.pi - np.pi * (np.arange(M) + 0.5) / M          # t_j = cos(theta_j)
    n = np.arange(N)
    if deriv == 0:
        return np.cos(np.outer(theta, n))
    if deriv == 1:
        # T_n'(cos theta) = n sin(n theta) / sin(theta); nodes avoid theta = 0, pi
        return (np.sin(np.outer(theta, n)) * n / np.sin(theta)[:, None]) * (2.0 / length)
    raise ValueError("deriv must be 0 or 1")


def cheb_forward_matrix(N: int, M: int) -> np.ndarray:
    """(N, M): c_n = (2 - delta_n0)/M sum_j f_j T_n(t_j), the Gauss-Chebyshev transform, which is
    the reference's (-1)^n DCT-II(f)/M with c_0 halved (transforms.py:117-121)."""
    theta = np.pi - np.pi * (np.arange(M) + 0.5) / M
    F = np.cos(np.outer(np.arange(N), theta)) * (2.0 / M)
    F[0] *= 0.5
    return F


def fourier_grid(M: int, a: float, Lx: float) -> np.ndarray:
    return a + Lx * np.arange(M) / M


def fourier_backward_matrix(K: int, M: int, deriv: int = 0, Lx: float = 2 * np.pi) -> np.ndarray:
    """(M, K) complex: f(x_j) = Re sum_k E[j, k] u_k with E = w_k (i k kappa)^deriv e^{i k kappa x_j},
    w_0 = 1, w_k = 2."""
    kappa = 2 * np.pi / Lx
    k = np.arange(K)
    w = np.full(K, 2.0)
    w[0] = 1.0
    E = np.exp(1j * np.outer(2 * np.pi * np.arange(M) / M, k)) * w
    if deriv:
        E = E * (1j * k * kappa) ** deriv
    return E


def fourier_forward_matrix(K: int, M: int) -> np.ndarray:
    """(K, M) complex: u_k = (1/M) sum_j f_j e^{-i k 2 pi j / M}."""
    return np.exp(-1j * np.outer(np.arange(K), 2 * np.pi * np.arange(M) / M)) / M


def _phase(kk: np.ndarray, M: int) -> np.ndarray:
    """2 pi (k m mod M) / M for every k of ``kk`` (rows) and grid index m < M (columns): the
    product is reduced exactly in integers before scaling, so the angles carry no large-argument
    rounding."""
    return 2 * np.pi * (np.outer(np.asarray(kk, dtype=np.int64), np.arange(M, dtype=np.int64)) % M) / M


def realfourier_forward_matrix(N: int, M: int) -> np.ndarray:
    """(N, M) real: the RealFourier coefficients [cos_0, msin_0, cos_1, msin_1, ...] of M grid values
    (reference/transforms.py:124-130: c_0 = F_0, c_2k = 2 Re F_k, c_2k+1 = 2 Im F_k with
    F = rfft / M; msin_0 = 0, no Nyquist mode)."""
    K = N // 2
    A = np.zeros((N, M))
    A[0] = 1.0 / M
    if K > 1:
        th = _phase(np.arange(1, K), M)
        A[2:N:2] = 2.0 * np.cos(th) / M
        A[3:N:2] = -2.0 * np.sin(th) / M
    return A


def realfourier_backward_matrix(N: int, M: int) -> np.ndarray:
    """(M, N) real: f_m = c_0 + sum_{k=1}^{N/2-1} (c_2k cos - c_2k+1 sin)(2 pi k m / M)
    (reference/transforms.py:170-177; c_1 = msin_0 does not enter)."""
    K = N // 2
    A = np.zeros((M, N))
    A[:, 0] = 1.0
    if K > 1:
        th = _phase(np.arange(1, K), M).T
        A[:, 2:N:2] = np.cos(th)
        A[:, 3:N:2] = -np.sin(th)
    return A


def complexfourier_freqs(N: int) -> np.ndarray:
    """Wavenumber of each ComplexFourier coefficient (fftfreq order; the Nyquist slot N/2 gets 0 and
    carries no mode)."""
    K = N // 2
    q = np.arange(N)
    f = np.where(q < K, q, q - N)
    f[K] = 0
    return f


def complexfourier_forward_matrix(N: int, M: int) -> np.ndarray:
    """(N, M) complex: c_q = (1/M) sum_m f_m e^{-2 pi i k_q m / M} for the kept wavenumbers k_q, the
    Nyquist row zero (reference/transforms.py:140-145)."""
    th = _phase(complexfourier_freqs(N) % M, M)
    A = np.exp(-1j * th) / M
    A[N // 2] = 0.0
    return A


def complexfourier_backward_matrix(N: int, M: int) -> np.ndarray:
    """(M, N) complex: f_m = sum_q c_q e^{2 pi i k_q m / M}, the Nyquist column zero
    (reference/transforms.py:181-187)."""
    th = _phase(complexfourier_freqs(N) % M, M).T
    A = np.exp(1j * th)
    A[:, N // 2] = 0.0
    return A


def integration_weights(N: int, length: float) -> np.ndarray:
    """int_a^b T_n(t(z)) dz = (L/2) * 2/(1 - n^2) for even n, 0 for odd n."""
    n = np.arange(N)
    w = np.zeros(N)
    even = n % 2 == 0
    w[even] = 2.0 / (1.0 - n[even] ** 2)
    return w * length / 2.0
