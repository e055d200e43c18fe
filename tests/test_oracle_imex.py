"""Pins of the IMEX oracle (oracle/imex.py): observed temporal orders on the manufactured heat
equation (as tests/core/test_core_imex.py:57-87), exact scheme constants, factor counts,
continuation rule."""
import numpy as np
import pytest

from oracle import spectral as S
from oracle.imex import ImexIntegrator, rk_tableau, sbdf_coefficients
from tests.imex_cases import N, exact, heat_system, make_F


def error(scheme, nsteps, T=0.5, phi="cos"):
    M, L, border = heat_system()
    integ = ImexIntegrator(M, L, border, scheme, T / nsteps, F=make_F(phi))
    x0 = np.zeros(N + 2)
    x0[:N] = S.cheb_forward(exact(0.0, phi), N)
    x, t = integ.run(x0, 0.0, nsteps)
    assert abs(t - T) < 1e-12
    return np.max(abs(S.cheb_backward(x[:N], N) - exact(T, phi))), integ


@pytest.mark.parametrize("scheme,order,phi", [
    ("SBDF1", 1, "cos"), ("SBDF2", 2, "cos"), ("SBDF3", 3, "flat"), ("SBDF4", 4, "flat"),
    ("RK111", 1, "cos"), ("RK222", 2, "cos"), ("RK443", 3, "cos"), ("RKSMR", 2, "cos"), ("RKGFY", 2, "cos"),
    ("CNAB1", 1, "cos"), ("CNAB2", 2, "cos"), ("MCNAB2", 2, "cos"), ("CNLF2", 2, "cos")])
def test_observed_order(scheme, order, phi):
    e1, _ = error(scheme, 40, phi=phi)
    e2, _ = error(scheme, 80, phi=phi)
    p = np.log2(e1 / e2)
    assert abs(p - order) < 0.1, (scheme, e1, e2, p)


def test_scheme_constants_are_consistent():
    # SBDF_q: sum a_j = 0 (constants are steady), sum c_j = 1 (F consistent), order conditions
    for q in (1, 2, 3, 4):
        a, c = sbdf_coefficients(q)
        assert abs(sum(a)) < 1e-15 and abs(sum(c) - 1) < 1e-15
        for p in range(1, q + 1):      # exact on t^p: sum_j a_j (-j)^p = p sum_j c_j (-j)^(p-1)
            lhs = sum(a[j] * (-j) ** p for j in range(q + 1))
            rhs = p * sum(c[j] * (-j) ** (p - 1) for j in range(q + 1))
            assert abs(lhs - rhs) < 1e-13, (q, p)
    for name in ("RK111", "RK222", "RK443"):
        A, H, c = rk_tableau(name)
        for i in range(len(c)):
            assert abs(sum(A[i]) - c[i]) < 1e-15 and abs(sum(H[i]) - c[i]) < 1e-15


def test_factor_counts_and_determinism():
    e, integ = error("RK443", 20)
    assert integ.n_factorizations == 1
    e2, _ = error("RK443", 20)
    assert e == e2
    _, sb = error("SBDF4", 20)
    assert sb.n_factorizations == 4


def test_continuation_rule():
    M, L, border = heat_system()
    integ = ImexIntegrator(M, L, border, "SBDF2", 0.01, F=make_F())
    x0 = np.zeros(N + 2)
    x, t = integ.run(x0, 0.0, 3)
    integ.run(x, t, 2)                              # continues
    with pytest.raises(ValueError):
        integ.run(x0, 0.0, 1)
    integ.reset()
    integ.run(x0, 0.0, 1)
