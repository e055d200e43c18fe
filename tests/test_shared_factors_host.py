"""Host checks of the sign-similarity test behind shared factors (conatus_amd/runtime/shared.py,
the reference's python/conatus/timestep/shared_factors.py:1-24 rule): B = D_r A D_c found exactly
when it holds, refused when any entry breaks it, when the patterns differ, for explicit zeros,
and when a border row would see its replaced diagonal change sign."""
import numpy as np
import scipy.sparse as sp

from conatus_amd.runtime.shared import sign_similarity
from tests.banded_cases import conditioned


def _pair(rng, n=40, r=2, border_ok=True):
    A = conditioned(rng, n, 3, 4, r).toarray()
    dr = rng.choice([-1.0, 1.0], n)
    dc = rng.choice([-1.0, 1.0], n)
    if border_ok:
        dc[:r] = dr[:r]
    else:
        dc[0] = -dr[0]
    return A, dr[:, None] * A * dc[None, :], dr, dc


def test_sign_similar_pairs_are_found_and_verified(rng):
    A, B, dr, dc = _pair(rng)
    got = sign_similarity(sp.csr_matrix(A), sp.csr_matrix(B), [0, 1])
    assert got is not None
    gr, gc = got
    assert np.array_equal(gr[:, None] * A * gc[None, :], B)
    assert np.all(gr[:2] * gc[:2] == 1)


def test_pairs_breaking_the_relation_are_refused(rng):
    A, B, _, _ = _pair(rng)
    C = B.copy()
    i, j = np.argwhere(A != 0)[7]
    C[i, j] *= 1.0 + 2.0 ** -52                     # one entry off by one ulp
    assert sign_similarity(sp.csr_matrix(A), sp.csr_matrix(C), [0, 1]) is None
    D = B.copy()
    D[i, j] = 0.0                                   # pattern differs
    assert sign_similarity(sp.csr_matrix(A), sp.csr_matrix(D), [0, 1]) is None
    E = sp.csr_matrix(A)
    E.data[3] = 0.0                                 # explicit zero
    assert sign_similarity(E, sp.csr_matrix(B), [0, 1]) is None
    # a sign pattern that does not factor as d_r[i] d_c[j]: flip one entry of B only
    F = B.copy()
    F[i, j] = -F[i, j]
    assert sign_similarity(sp.csr_matrix(A), sp.csr_matrix(F), [0, 1]) is None


def test_border_rows_must_keep_their_diagonal_sign(rng):
    A, B, _, _ = _pair(rng, border_ok=False)
    assert sign_similarity(sp.csr_matrix(A), sp.csr_matrix(B), []) is not None
    assert sign_similarity(sp.csr_matrix(A), sp.csr_matrix(B), [0, 1]) is None


def test_channel_pairs_fail_the_border_rule():
    """The primitive-variable channel's (kx, kz) and (kx, -kz) systems are sign-similar (the odd
    terms i kz flip with the w unknowns and rows), but a w boundary row would see its replaced
    diagonal change sign, so the reference's rule does not share them."""
    from conatus_amd.problems import Channel3D
    pb = Channel3D(Nx=8, Ny=13, Nz=8, Re=800.0)
    m1, m2 = 1 * pb.Kz + 1, 1 * pb.Kz + (pb.Kz - 1)
    A = 1000.0 * pb.mats[m1][0] + pb.mats[m1][1]
    B = 1000.0 * pb.mats[m2][0] + pb.mats[m2][1]
    assert sign_similarity(A, B, []) is not None
    assert sign_similarity(A, B, pb.border) is None
