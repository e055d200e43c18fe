"""CPU tests of the 3D channel discretization (conatus_amd.problems.Channel3D) and its oracle
(oracle/channel3d.py).

* Operators: the reference's manufactured Stokes-resolvent mode (tests/channel3d/stokes_mode.py,
  tests/channel3d/test_influence.py:50-63): u = (i kx, ., i kz) v'/k^2, v = cos^2(pi y/2),
  p = sin(pi y/2) + y^2 on the mode (kx, kz) = (1, 2), c = 50, Re = 100, solved with c M + L of
  that mode.  The reference's generalized tau reaches 6.8e-11 on the N = 16 grid and < 1e-13 at
  N = 24; a wrong sign, a missing term or a misplaced tau fails both.
* Explicit term: against the exact Fourier-Chebyshev projection of -u.grad(u) computed by brute
  force (convolution over the modes, numpy.polynomial.chebyshev products in y) for a band-limited
  field, which the 3/2 grid resolves exactly.
* Laminar Poiseuille flow is a steady state; the initial perturbation is divergence-free and
  vanishes at the walls."""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from conatus_amd import spectral as S
from conatus_amd.problems import P, UX, UY, UZ, Channel3D
from oracle.banded import BorderedBanded
from oracle.channel3d import Channel3DOracle


def _stokes_exact(kx, kz, c, Re):
    k2, pi = kx * kx + kz * kz, np.pi
    v1 = lambda y: -(pi / 2) * np.sin(pi * y)                                    # noqa: E731
    v3 = lambda y: (pi ** 3 / 2) * np.sin(pi * y)                                # noqa: E731
    vel = (lambda y: 1j * kx * v1(y) / k2, lambda y: (1 + np.cos(pi * y)) / 2, lambda y: 1j * kz * v1(y) / k2)
    d2 = (lambda y: 1j * kx * v3(y) / k2, lambda y: -(pi ** 2 / 2) * np.cos(pi * y), lambda y: 1j * kz * v3(y) / k2)
    p = lambda y: np.sin(pi * y / 2) + y ** 2                                    # noqa: E731
    gp = (lambda y: 1j * kx * p(y), lambda y: (pi / 2) * np.cos(pi * y / 2) + 2 * y, lambda y: 1j * kz * p(y))
    r = [lambda y, i=i: c * vel[i](y) - (d2[i](y) - k2 * vel[i](y)) / Re + gp[i](y) for i in range(3)]
    return vel, p, r


def _coeffs(f, N):
    re = C.chebinterpolate(lambda t: np.real(f(t) + 0j * t), N - 1)
    im = C.chebinterpolate(lambda t: np.imag(f(t) + 0j * t), N - 1)
    return re + 1j * im


def _stokes_error(N, points):
    KX, KZ, CC, RE = 1.0, 2.0, 50.0, 100.0
    pb = Channel3D(Nx=4, Ny=N, Nz=4, Re=RE, Lx=2 * np.pi / KX, Lz=2 * np.pi / KZ)
    M, L = pb.mats[pb.mode(1, 1)]
    A = (CC * M + L).tocsr()
    vel, p, r = _stokes_exact(KX, KZ, CC, RE)
    b = np.zeros(pb.n, complex)
    for e in range(3):
        b[pb.equation_rows(e)] = pb.S02 @ _coeffs(r[e], N)
    x = BorderedBanded(A, pb.border, np.complex128).solve(b)
    yy = points(N)
    err_u = max(np.abs(C.chebval(yy, x[pb.field_index(f)]) - vel[f](yy)).max() for f in (UX, UY, UZ))
    err_p = np.abs(C.chebval(yy, x[pb.field_index(P)]) - p(yy)).max()
    return err_u, err_p, np.abs(A @ x - b).max()


def test_stokes_resolvent_mode_matches_the_reference_tau():
    grid = lambda N: S.cheb_grid(N, -1.0, 1.0)                                   # noqa: E731
    e16 = _stokes_error(16, grid)
    e24 = _stokes_error(24, grid)
    assert 5e-11 < e16[0] < 9e-11          # the reference's tau: 6.8e-11 on the N = 16 grid
    assert e24[0] < 1e-13 and e24[1] < 1e-12
    assert e16[2] < 1e-13 and e24[2] < 1e-13


def _full(V, pb):
    """Real field of (Kx, Kz, N) modes -> {(kx, kz): Chebyshev coefficients} over kx of both signs."""
    out = {}
    for kx in range(pb.Kx):
        for iz in range(pb.Kz):
            kz = int(pb.kz_of(iz))
            c = V[kx, iz]
            if not np.any(c):
                continue
            out[(kx, kz)] = out.get((kx, kz), 0) + c
            if kx > 0:
                out[(-kx, -kz)] = out.get((-kx, -kz), 0) + np.conj(c)
    return out


def _brute_explicit(pb, X):
    V = pb.velocity(X)
    comps = [_full(V[f], pb) for f in range(3)]

    def deriv(d, axis):
        out = {}
        for (kx, kz), c in d.items():
            if axis == 0:
                out[(kx, kz)] = 1j * kx * pb.kx0 * c
            elif axis == 2:
                out[(kx, kz)] = 1j * kz * pb.kz0 * c
            else:
                out[(kx, kz)] = C.chebder(c) if len(c) > 1 else 0 * c
        return out

    F = pb.b0().copy()
    for e in range(3):
        acc = {}
        for a in range(3):                   # -(u_a d_a) u_e
            da = deriv(comps[e], a)
            for k1, c1 in comps[a].items():
                for k2, c2 in da.items():
                    k = (k1[0] + k2[0], k1[1] + k2[1])
                    prod = np.zeros(2 * pb.N, complex)
                    m = C.chebmul(c1, c2)
                    prod[:m.size] = m
                    acc[k] = acc.get(k, 0) - prod
        c = np.zeros((pb.Kx, pb.Kz, pb.N), complex)
        for (kx, kz), v in acc.items():
            if 0 <= kx < pb.Kx and abs(kz) < pb.Nz // 2:
                v = np.asarray(v)[:pb.N]
                c[kx, pb.mode(0, kz) % pb.Kz, :v.size] = v
        F[:, pb.equation_rows(e)] += (pb.S02 @ c.reshape(pb.K, pb.N).T).T
    return F


@pytest.mark.parametrize("shape", [(8, 9, 8), (6, 12, 10)])
def test_explicit_term_equals_brute_force_projection(shape):
    nx, ny, nz = shape
    pb = Channel3D(Nx=nx, Ny=ny, Nz=nz, Re=300.0, Lx=3.0, Lz=2.0)
    X = pb.initial_state(seed=5, energy=1e-2, kmax=1)
    got = Channel3DOracle(pb).explicit(X)
    ref = _brute_explicit(pb, X)
    rows = np.concatenate([pb.equation_rows(e) for e in range(3)])
    scale = np.abs(ref[:, rows]).max()
    assert scale > 1e-3
    assert np.abs(got[:, rows] - ref[:, rows]).max() < 1e-13 * max(1.0, scale)


def test_poiseuille_is_steady_and_perturbation_admissible():
    pb = Channel3D(Nx=8, Ny=13, Nz=8, Re=200.0)
    o = Channel3DOracle(pb, "SBDF2", 1e-2)
    X0 = pb.poiseuille()
    assert np.array_equal(o.explicit(X0), pb.b0())
    X = o.run(X0, 5)
    assert np.abs(X - X0)[:, pb.field_index(UX)].max() < 1e-12
    Xp = pb.initial_state(seed=2, energy=1e-4) - X0
    V = pb.velocity(Xp)
    e = pb.kinetic_energy(V)
    assert abs(e - 1e-4) < 1e-12
    ikx = (1j * pb.kx0 * np.arange(pb.Kx))[:, None, None]
    ikz = (1j * pb.kz0 * pb.kz_of(np.arange(pb.Kz)))[None, :, None]
    div = ikx * V[0] + np.apply_along_axis(lambda c: np.append(C.chebder(c), 0), 2, V[1]) + ikz * V[2]
    assert np.abs(div).max() < 1e-14
    for f in range(3):
        for t in (-1.0, 1.0):
            assert np.abs(C.chebval(t, np.moveaxis(V[f], 2, 0))).max() < 1e-15
    # conjugate pairs at kx = 0 (real fields)
    for kz in range(1, pb.Nz // 2):
        assert np.array_equal(V[:, 0, pb.mode(0, -kz)], np.conj(V[:, 0, pb.mode(0, kz)]))


def test_every_mode_steps_with_the_oracle():
    pb = Channel3D(Nx=6, Ny=10, Nz=6, Re=500.0)
    o = Channel3DOracle(pb, "SBDF3", 5e-3)
    X = o.run(pb.initial_state(seed=1, energy=1e-3), 4)
    assert np.all(np.isfinite(X))
    d = pb.velocity(X - pb.poiseuille())
    assert 0 < pb.kinetic_energy(d) < 2e-3


def test_rounding_floor_of_a_trajectory():
    """Two backward-stable solves differ, over 12 SBDF3 steps, by more than 1e-11 in the velocity:
    the oracle against itself with every factored matrix perturbed by half an ulp (the bound of
    the device parity tests of the trajectories rests on this floor)."""
    from oracle.banded import BorderedBanded

    class Perturbed(Channel3DOracle):
        rng = np.random.default_rng(1)

        def factor(self, k, alpha, beta=1.0):
            key = (k, alpha, beta)
            if key not in self.factors:
                A = (np.complex128(alpha) * self.M[k] + np.complex128(beta) * self.L[k]).tocsr()
                A.data = A.data * (1 + 1e-16 * self.rng.standard_normal(A.data.shape))
                self.factors[key] = BorderedBanded(A, self.pb.border, np.complex128)
            return self.factors[key]

    pb = Channel3D(Nx=8, Ny=13, Nz=8, Re=500.0)
    X0 = pb.initial_state(seed=4, energy=1e-3, kmax=2)
    ref = Channel3DOracle(pb, "SBDF3", 2e-3).run(X0, 12)
    got = Perturbed(pb, "SBDF3", 2e-3).run(X0, 12)
    vel = np.concatenate([pb.field_index(f) for f in (UX, UY, UZ)])
    floor = np.abs(got - ref)[:, vel].max() / np.abs(ref[:, vel]).max()
    assert 1e-11 < floor < 2e-10
