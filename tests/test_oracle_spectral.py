"""Closed-form pins of oracle/spectral.py (transforms and ultraspherical operators)."""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest
from scipy.special import eval_gegenbauer

from oracle import spectral as S


@pytest.mark.parametrize("M,N", [(24, 16), (48, 32), (145, 97)])
def test_chebyshev_forward_of_nodes_is_identity(M, N):
    t = S.cheb_grid(M)
    for n in (0, 1, 5, N - 1):
        f = np.cos(n * np.arccos(t))               # T_n on the Gauss-Chebyshev nodes
        c = S.cheb_forward(f, N)
        e = np.zeros(N)
        e[n] = 1.0
        assert np.max(abs(c - e)) < 1e-13


def test_chebyshev_round_trip_and_grid_order():
    rng = np.random.default_rng(0)
    c = rng.standard_normal(40)
    f = S.cheb_backward(c, 60)
    assert np.max(abs(S.cheb_forward(f, 40) - c)) < 1e-13
    assert np.all(np.diff(S.cheb_grid(60)) > 0)
    assert np.max(abs(f - C.chebval(S.cheb_grid(60), c))) < 1e-12


def test_realfourier_pairs():
    M, N, L = 48, 32, 2 * np.pi * 1.5
    x = S.fourier_grid(M, 0.3, L)
    k = 2 * np.pi / L
    f = 2.0 + 3.0 * np.cos(2 * k * (x - 0.3)) - 5.0 * np.sin(7 * k * (x - 0.3))
    c = S.rf_forward(f, N)
    e = np.zeros(N)
    e[0], e[4], e[15] = 2.0, 3.0, 5.0             # (-sin) coefficient of mode 7 at 2*7+1
    assert np.max(abs(c - e)) < 1e-13
    assert np.max(abs(S.rf_backward(c, M) - f)) < 1e-12


def test_differentiation_and_conversion_closed_forms():
    N = 20
    rng = np.random.default_rng(1)
    c = rng.standard_normal(N)
    c[-2:] = 0                                     # so that nothing leaves the space
    t = np.linspace(-0.9, 0.9, 7)
    # D: T -> C^(1) and S_0: T -> C^(1): S_0^-1 D c equals chebder(c)
    d1 = S.differentiation(N, 0) @ c
    der = C.chebder(c)
    val_der = C.chebval(t, der)
    val_d1 = sum(d1[n] * eval_gegenbauer(n, 1, t) for n in range(N))
    assert np.max(abs(val_d1 - val_der)) < 1e-11
    # S_lam: C^(lam) -> C^(lam+1) preserves the function
    for lam in (0, 1, 2):
        a = rng.standard_normal(N)
        a[-2:] = 0
        b = S.conversion(N, lam) @ a
        if lam == 0:
            va = C.chebval(t, a)
        else:
            va = sum(a[n] * eval_gegenbauer(n, lam, t) for n in range(N))
        vb = sum(b[n] * eval_gegenbauer(n, lam + 1, t) for n in range(N))
        assert np.max(abs(va - vb)) < 1e-12
    # second derivative T -> C^(2)
    d2 = S.differentiation(N, 0, 2) @ c
    v2 = sum(d2[n] * eval_gegenbauer(n, 2, t) for n in range(N))
    assert np.max(abs(v2 - C.chebval(t, C.chebder(c, 2)))) < 1e-10


def test_cheb_diff_recurrence_matches_chebder():
    rng = np.random.default_rng(2)
    c = rng.standard_normal(33)
    d = S.cheb_diff(c, L=2.0)
    ref = np.zeros(33)
    ref[:32] = C.chebder(c)
    assert np.max(abs(d - ref)) < 1e-12
    d4 = S.cheb_diff(c, L=4.0)
    assert np.max(abs(d4 - ref / 2)) < 1e-12


def test_complexfourier_modes_and_round_trip():
    """Closed form: f = 2 + (1 - 3i) e^{i 3 k x} + 4i e^{-i 5 k x} has coefficients 2 at wavenumber
    0, 1 - 3i at +3 and 4i at -5 (fftfreq order: index N - 5), and nothing in the Nyquist slot."""
    M, N, L = 48, 32, 3.0
    x = S.fourier_grid(M, -0.7, L)
    k = 2 * np.pi / L
    f = 2.0 + (1 - 3j) * np.exp(1j * 3 * k * (x + 0.7)) + 4j * np.exp(-1j * 5 * k * (x + 0.7))
    c = S.cf_forward(f, N)
    e = np.zeros(N, complex)
    e[0], e[3], e[N - 5] = 2.0, 1 - 3j, 4j
    assert np.max(abs(c - e)) < 1e-13 and c[N // 2] == 0
    assert np.max(abs(S.cf_backward(c, M) - f)) < 1e-12
    rng = np.random.default_rng(1)
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    c[N // 2] = 0
    assert np.max(abs(S.cf_forward(S.cf_backward(c, M), N) - c)) < 1e-13


@pytest.mark.parametrize("M,N", [(12, 8), (48, 32), (145, 97), (192, 128)])
def test_product_transform_matrices_are_the_oracle_transforms(M, N):
    """conatus_amd.spectral's dense transform matrices (what the device applies) against the
    oracle's FFT / DCT transforms, on random data."""
    from conatus_amd import spectral as P
    rng = np.random.default_rng(M)
    f = rng.standard_normal((M, 3))
    c = rng.standard_normal((N, 3))
    assert np.max(abs(P.cheb_forward_matrix(N, M) @ f - S.cheb_forward(f, N))) < 1e-13
    assert np.max(abs(P.cheb_backward_matrix(N, M) @ c - S.cheb_backward(c, M))) < 1e-12
    if N % 2 == 0:
        assert np.max(abs(P.realfourier_forward_matrix(N, M) @ f - S.rf_forward(f, N))) < 1e-13
        cr = c.copy()
        cr[1] = 0.0
        assert np.max(abs(P.realfourier_backward_matrix(N, M) @ cr - S.rf_backward(cr, M))) < 1e-12
        fz = f + 1j * rng.standard_normal((M, 3))
        cz = c + 1j * rng.standard_normal((N, 3))
        cz[N // 2] = 0
        assert np.max(abs(P.complexfourier_forward_matrix(N, M) @ fz - S.cf_forward(fz, N))) < 1e-13
        assert np.max(abs(P.complexfourier_backward_matrix(N, M) @ cz - S.cf_backward(cz, M))) < 1e-12
