"""CPU oracle of the spectral conventions — TEST INFRASTRUCTURE ONLY (see oracle/banded.py).

Restates, in numpy / scipy.fft:
* Olver–Townsend operators on [-1, 1] (python/conatus/discretize/ultraspherical.py:43-90):
  conversion S_lam: C^(lam) -> C^(lam+1) (T -> C^(1): diag (1, 1/2, 1/2, ...), +2 diagonal -1/2;
  lam >= 1: diag lam/(n+lam), +2 diagonal -lam/(n+lam) at n >= 2), conversion chains,
  differentiation (T -> C^(1): n on the +1 diagonal; C^(k) -> C^(k+1): 2k), all N x N;
* grid <-> coefficient transforms (python/conatus/reference/transforms.py:86-177):
  Chebyshev on ascending Gauss-Chebyshev nodes t_j = -cos(pi (j + 1/2) / M), forward
  c = (-1)^n DCT-II(f)/M with c_0 /= 2 (first N kept), backward by the inverse;
  RealFourier pairs [cos_0, msin_0, cos_1, msin_1, ...] (Nyquist not represented),
  f = sum c_k cos(k (x-a) 2pi/L) + s_k (-sin(k (x-a) 2pi/L));
  ComplexFourier (transforms.py:140-145, 181-187): fft / M in fftfreq order, Nyquist slot zero.
* the Chebyshev derivative backward recurrence (python/conatus/timestep/spectral.py:53-82).

Pinned by tests/test_oracle_spectral.py with closed forms: T_n(t_j) = cos(n acos t_j), the
derivative of known polynomials (numpy.polynomial.chebyshev.chebder), C^(lam) conversion checked
by evaluating both sides with scipy.special.eval_gegenbauer, Fourier modes of cos/sin and of
complex exponentials of positive and negative wavenumber.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp


def conversion(N: int, lam: int) -> sp.csr_matrix:
    n = np.arange(N, dtype=float)
    if lam == 0:
        diag = np.full(N, 0.5)
        diag[0] = 1.0
        upper = -0.5 * np.ones(max(N - 2, 0))
    else:
        diag = lam / (n + lam)
        upper = -lam / (n[2:] + lam)
    return sp.diags([diag, upper], [0, 2], shape=(N, N), format="csr")


def conversion_chain(N: int, lam_from: int, lam_to: int) -> sp.csr_matrix:
    out = sp.identity(N, format="csr")
    for lam in range(lam_from, lam_to):
        out = conversion(N, lam) @ out
    return out.tocsr()


def differentiation(N: int, lam: int, order: int = 1) -> sp.csr_matrix:
    out = sp.identity(N, format="csr")
    for k in range(lam, lam + order):
        n = np.arange(1, N, dtype=float)
        vals = n if k == 0 else np.full(N - 1, 2.0 * k)
        out = sp.diags([vals], [1], shape=(N, N), format="csr") @ out
    return out.tocsr()


def cheb_grid(M: int, a: float = -1.0, b: float = 1.0) -> np.ndarray:
    t = -np.cos(np.pi * (np.arange(M) + 0.5) / M)
    return a + (t + 1) * (b - a) / 2


def cheb_forward(f: np.ndarray, N: int, axis: int = 0) -> np.ndarray:
    f = np.moveaxis(np.asarray(f), axis, 0)
    M = f.shape[0]
    y = sfft.dct(f, type=2, axis=0) / M
    y[0] /= 2
    sign = ((-1.0) ** np.arange(M)).reshape((M,) + (1,) * (f.ndim - 1))
    return np.moveaxis((y * sign)[:N], 0, axis)


def cheb_backward(c: np.ndarray, M: int, axis: int = 0) -> np.ndarray:
    c = np.moveaxis(np.asarray(c), axis, 0)
    N = c.shape[0]
    y = np.zeros((M,) + c.shape[1:], dtype=c.dtype)
    sign = ((-1.0) ** np.arange(N)).reshape((N,) + (1,) * (c.ndim - 1))
    y[:N] = c * sign
    y[0] *= 2
    y *= M
    return np.moveaxis(sfft.idct(y, type=2, axis=0), 0, axis)


def fourier_grid(M: int, a: float, L: float) -> np.ndarray:
    return a + L * np.arange(M) / M


def rf_forward(f: np.ndarray, N: int, axis: int = 0) -> np.ndarray:
    f = np.moveaxis(np.asarray(f), axis, 0)
    M = f.shape[0]
    F = np.fft.rfft(f, axis=0) / M
    K = N // 2
    c = np.zeros((N,) + f.shape[1:])
    c[0] = F[0].real
    c[2:N:2] = 2 * F[1:K].real
    c[3:N:2] = 2 * F[1:K].imag
    return np.moveaxis(c, 0, axis)


def rf_backward(c: np.ndarray, M: int, axis: int = 0) -> np.ndarray:
    c = np.moveaxis(np.asarray(c), axis, 0)
    N = c.shape[0]
    K = N // 2
    F = np.zeros((M // 2 + 1,) + c.shape[1:], dtype=complex)
    F[0] = c[0]
    F[1:K] = (c[2:N:2] + 1j * c[3:N:2]) / 2
    F *= M
    return np.moveaxis(np.fft.irfft(F, n=M, axis=0), 0, axis)


def cf_forward(f: np.ndarray, N: int, axis: int = 0) -> np.ndarray:
    """ComplexFourier forward (reference/transforms.py:140-145): fft / M, the first N/2 wavenumbers
    then the last N/2 - 1 (fftfreq order), the Nyquist slot N/2 zero."""
    f = np.moveaxis(np.asarray(f), axis, 0)
    M = f.shape[0]
    F = np.fft.fft(f, axis=0) / M
    K = N // 2
    c = np.zeros((N,) + f.shape[1:], dtype=complex)
    c[:K] = F[:K]
    c[K + 1:] = F[M - K + 1:]
    return np.moveaxis(c, 0, axis)


def cf_backward(c: np.ndarray, M: int, axis: int = 0) -> np.ndarray:
    """ComplexFourier backward (reference/transforms.py:181-187): the spectrum zero padded to M
    (Nyquist slot dropped), times M, inverse fft."""
    c = np.moveaxis(np.asarray(c), axis, 0)
    N = c.shape[0]
    K = N // 2
    F = np.zeros((M,) + c.shape[1:], dtype=complex)
    F[:K] = c[:K]
    F[M - K + 1:] = c[K + 1:]
    return np.moveaxis(np.fft.ifft(F * M, axis=0), 0, axis)


def cheb_diff(c: np.ndarray, L: float = 2.0, axis: int = 0) -> np.ndarray:
    """d/dx of a Chebyshev series on an interval of length L: out[n] = out[n+2] + 2(n+1) c[n+1],
    out[0] /= 2, times 2/L (timestep/spectral.py:53-82)."""
    c = np.moveaxis(np.asarray(c), axis, 0)
    N = c.shape[0]
    out = np.zeros_like(c)
    if N > 1:
        out[N - 2] = 2 * (N - 1) * c[N - 1]
    for n in range(N - 3, -1, -1):
        out[n] = out[n + 2] + 2 * (n + 1) * c[n + 1]
    out[0] = out[0] / 2
    return np.moveaxis(out * (2.0 / L), 0, axis)
