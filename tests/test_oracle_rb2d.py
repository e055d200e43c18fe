"""Pins of the RB 2D discretization and of the oracle's explicit term (CPU)."""
import numpy as np
import pytest
import numpy.polynomial.chebyshev as C
import scipy.linalg as la
import scipy.sparse as sp

from conatus_amd.problems import TEQ, W, RayleighBenard2D
from oracle.rb2d import RB2DOracle, from_pairs, to_pairs


def _growth(Ra, kx, Nz=32, Pr=1.0):
    pb = RayleighBenard2D(Nx=4, Nz=Nz, Lx=2 * np.pi / kx, Ra=Ra, Pr=Pr, z0=-0.5, z1=0.5)
    M, L = pb.mats[1]
    Wm = sp.lil_matrix(L.shape, dtype=complex)            # linearized -u.grad(T_cond) = +w / Lz
    B = sp.coo_matrix(pb.S02)
    for r_, c_, v in zip(B.row, B.col, B.data):
        Wm[pb.row(r_, TEQ), pb.col(c_, W)] = v / pb.Lz
    ev = la.eig(-(L - Wm.tocsr()).toarray(), M.toarray(), right=False)
    ev = ev[np.isfinite(ev)]
    return ev[np.argmax(ev.real)]


def test_linear_growth_rate_matches_reference_validation():
    """sigma = 5.5425059 at Ra = 2500, k = pi, Pr = 1 (python/conatus/problems/convection.py:38-40)
    and neutral stability at the Reid-Harris point Ra_c = 1707.762, k_c = 3.117."""
    s = _growth(2500.0, np.pi)
    assert abs(s.real - 5.5425059) < 5e-7 and abs(s.imag) < 1e-9
    s0 = _growth(1707.762, 3.117)
    assert abs(s0.real) < 1e-4


def test_mode_systems_are_square_banded_and_uniform():
    """Every mode: square n = 4 Nz + 7, the banded part B of the Woodbury split invertible, and the
    common band kl = 12, ku = 18 (the band the reference reports for the RB 2D k = 0 system)."""
    from oracle.banded import BorderedBanded
    pb = RayleighBenard2D(Nx=16, Nz=12, Lx=4.0, Ra=1e5, Pr=1.0)
    assert len(pb.mats) == 8
    kls, kus = [], []
    for M, L in pb.mats:
        assert M.shape == L.shape == (pb.n, pb.n)
        o = BorderedBanded((1.5e3 * M + L).tocsr(), pb.border, np.complex128)   # raises if B singular
        kls.append(o.kl)
        kus.append(o.ku)
    assert max(kls) == 12 and max(kus) == 18


def test_pairs_round_trip():
    rng = np.random.default_rng(0)
    c = rng.standard_normal((8, 5)) + 1j * rng.standard_normal((8, 5))
    c[0] = c[0].real
    assert np.allclose(from_pairs(to_pairs(c)), c)


def test_explicit_term_closed_form():
    """u = 2 Re(a e^{i kappa x}) f(z), w = 0, T = 2 Re(b e^{i kappa x}) g(z): -u T_x has mode 0
    coefficient -Re(a conj(i kappa b)) f g * 2 / 2 ... computed exactly with chebmul; the product
    is representable, so the dealiased evaluation is exact to rounding."""
    pb = RayleighBenard2D(Nx=8, Nz=10, Lx=2 * np.pi, Ra=1.0, Pr=1.0)
    kap = pb.kappa
    a, b = 0.7 - 0.2j, -0.3 + 0.5j
    f = np.array([0.5, -0.25, 0.1, 0.0])          # Chebyshev coefficients (degree 2)
    g = np.array([0.2, 0.3, -0.1, 0.05])          # degree 3
    X = np.zeros((pb.K, pb.n), complex)
    X[1, pb.field_index(0)[:4]] = a * f           # u_hat_1 = a f
    X[1, pb.field_index(2)[:4]] = b * g           # T_hat_1 = b g
    F = RB2DOracle(pb).explicit(X)
    fg = C.chebmul(f, g)                          # degree 5 < N
    # u = 2 Re(a f e^{i t}), T_x = 2 Re(i kap b g e^{i t});  product modes:
    #   mode 0: 2 Re(a f conj(i kap b g)) ; mode 2: 2 (a f)(i kap b g) / ... in u_hat convention
    m0 = -2 * np.real(a * np.conj(1j * kap * b)) * fg
    m2 = -(a * 1j * kap * b) * fg
    e0 = np.zeros(pb.N, complex)
    e0[:fg.size] = m0
    e2 = np.zeros(pb.N, complex)
    e2[:fg.size] = m2
    rows = pb.equation_rows(TEQ)
    assert np.max(abs(F[0, rows] - pb.S02 @ e0)) < 1e-14
    assert np.max(abs(F[2, rows] - pb.S02 @ e2)) < 1e-14
    assert np.max(abs(F[1, rows])) < 1e-14 and np.max(abs(F[3, rows])) < 1e-14


def _brute_explicit_rb2d(pb, X):
    """-(u d_x + w d_z) of (u, w, T) by explicit Fourier-Chebyshev convolution: every pair of
    modes (k1, k2) of both signs multiplied with chebmul, summed into mode k1 + k2, truncated to
    |k| < K and n < N, converted by S_{0->2} (no grid, no transform, no dealiasing rule)."""
    K, N = pb.K, pb.N
    cols = [pb.field_index(f) for f in range(3)]
    full = []                                     # {k: coefficients} over k = -(K-1) .. K-1
    for f in range(3):
        c = X[:, cols[f]]
        d = {k: c[k].copy() for k in range(K)}
        d.update({-k: np.conj(c[k]) for k in range(1, K)})
        full.append(d)
    dz = 2.0 / pb.Lz

    def ddx(d):
        return {k: 1j * k * pb.kappa * v for k, v in d.items()}

    def ddz(d):
        return {k: (C.chebder(v) * dz if len(v) > 1 else 0 * v) for k, v in d.items()}

    F = RB2DOracle(pb).b0.copy()
    u, w = full[0], full[1]
    for e, f in enumerate((0, 1, 2)):             # x momentum (u), z momentum (w), temperature (T)
        acc = {}
        for vel, der in ((u, ddx(full[f])), (w, ddz(full[f]))):
            for k1, c1 in vel.items():
                for k2, c2 in der.items():
                    k = k1 + k2
                    if 0 <= k < K:
                        m = C.chebmul(c1, c2)[:N]
                        v = np.zeros(N, complex)
                        v[:m.size] = m
                        acc[k] = acc.get(k, 0) - v
        c = np.zeros((K, N), complex)
        for k, v in acc.items():
            c[k] = v
        F[:, pb.equation_rows(e)] = (pb.S02 @ c.T).T
    return F


@pytest.mark.parametrize("shape", [(8, 10), (12, 9)])
def test_explicit_term_with_vertical_velocity_equals_brute_force_projection(shape):
    """u, w and T all non-zero and of low degree (every product representable on the 3/2 grid):
    the oracle's transform-based explicit term equals the brute-force convolution to rounding, so
    the -w d_z terms are pinned (a swapped u_z / w_z or a wrong d/dz scale fails)."""
    nx, nz = shape
    pb = RayleighBenard2D(Nx=nx, Nz=nz, Lx=3.0, Ra=1.0, Pr=1.0)
    rng = np.random.default_rng(nx * 31 + nz)
    X = np.zeros((pb.K, pb.n), complex)
    deg = nz // 2
    for f in range(3):
        c = (rng.standard_normal((pb.K, deg)) + 1j * rng.standard_normal((pb.K, deg))) * 0.3
        c[0] = c[0].real
        c[pb.K // 2 + 1:] = 0.0                   # modes up to K/2: products stay below K
        X[:, pb.field_index(f)[:deg]] = c
    got = RB2DOracle(pb).explicit(X)
    ref = _brute_explicit_rb2d(pb, X)
    rows = np.concatenate([pb.equation_rows(e) for e in range(3)])
    scale = np.abs(ref[:, rows]).max()
    assert scale > 1e-2
    assert np.abs(got[:, rows] - ref[:, rows]).max() < 1e-13 * scale
