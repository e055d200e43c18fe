"""GPU parity of the gfx950 bordered-banded factor/solve against the CPU oracle (oracle/banded.py).

Bar: normwise backward error <= 1e-13, the residual within
10x the oracle's own residual plus n eps (the reference's cross-backend rule,
tests/cross_backend/test_banded_level.py:71-76), solutions within a conditioning-scaled distance of
the oracle's, fused gather/scatter and split-complex maps exact, bitwise determinism.  Every test
runs in every solve mode of the library (CNHIP_MODE=lane: one lane per system and right-hand side
over 16-system interleaved factors; CNHIP_MODE=wave: one wavefront per system and right-hand side,
U stored by columns; CNHIP_MODE=part: wave mode with each system's sweeps partitioned over 2-8
wavefronts)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.banded import BorderedBanded as OracleBB
from tests.banded_cases import (almost_banded, backward_error, bordered_case, conditioned, csr_of,
                                microbench_csr, sparse_banded)

pytestmark = pytest.mark.gpu

BE_TOL = 1e-13


@pytest.fixture(autouse=True, params=["lane", "lane8", "wave", "part"])
def solve_mode(request, monkeypatch):
    """lane: streamed sweeps in blocks of 16 steps where they fit in LDS; lane8: blocks of 8;
    part: wave mode with each system's sweeps partitioned over 2-8 wavefronts where the system is
    long enough (else wave mode)."""
    mode = request.param.rstrip("8")
    monkeypatch.setenv("CNHIP_MODE", mode)
    if request.param == "lane8":
        monkeypatch.setenv("CNHIP_SPX", "8")
    print(f"solve mode {request.param}", flush=True)
    return mode


def _torch():
    import torch
    return torch


def _batch(rng, B, n, kl, ku, r, dtype):
    border = sorted(rng.choice(n, size=r, replace=False).tolist())
    order = border + [i for i in range(n) if i not in border]
    mats = []
    for _ in range(B):
        Ap = almost_banded(rng, n, kl, ku, range(r), dtype)
        A = np.empty_like(Ap)
        A[order] = Ap
        mats.append(sp.csr_matrix(A))
    return mats, border


def _rhs(rng, shape, dtype):
    b = rng.standard_normal(shape)
    if np.dtype(dtype) == np.complex128:
        b = b + 1j * rng.standard_normal(shape)
    return b


def _check_against_oracle(mats, border, b, x, dtype, well_posed=True):
    """well_posed: the oracle's own backward error meets the bar, so the GPU must too.  Otherwise
    (random non-dominant bands, where the Woodbury form of the reference is itself unstable) the
    reference's cross-backend rule applies: residual within 10x the oracle's plus n eps."""
    for k, A in enumerate(mats):
        try:
            o = OracleBB(A, border, dtype)
        except RuntimeError:
            continue                       # singular for the oracle: no reference solution
        xo = o.solve(b[k])
        res_o = np.max(abs(A @ xo - b[k]))
        res = np.max(abs(A @ x[k] - b[k]))
        n = A.shape[0]
        be, be_o = backward_error(A, x[k], b[k]), backward_error(A, xo, b[k])
        if well_posed:
            assert be_o <= BE_TOL, (k, be_o)
            assert be <= BE_TOL, (k, be)
            assert np.max(abs(x[k] - xo)) <= 1e-12 * max(1.0, o.cond) * np.max(abs(xo)) * n
        else:
            if be_o > 1e-12:
                continue                   # the Woodbury split itself fails on this matrix
            assert be <= max(BE_TOL, 10 * be_o), (k, be, be_o)
        assert res <= 10 * res_o + 1e-13 * n * np.max(abs(b[k])), (k, res, res_o)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("r", [0, 1, 2, 3, 4])
def test_batch_matches_oracle(dtype, r, rng):
    from conatus_amd.runtime import BorderedBandedBatch
    n, kl, ku = int(rng.integers(20, 80)), int(rng.integers(0, 4)), int(rng.integers(0, 4))
    print(f"case n={n} kl={kl} ku={ku} r={r} dtype={np.dtype(dtype).name}", flush=True)
    mats, border = _batch(rng, 7, n, kl, ku, r, dtype)
    bb = BorderedBandedBatch(mats, border)
    b = _rhs(rng, (7, n), dtype)
    x = bb.solve(b)
    assert x.shape == b.shape and x.dtype == np.dtype(dtype)
    _check_against_oracle(mats, border, b, x, dtype)


@pytest.mark.parametrize("kl,ku", [(1, 0), (2, 7), (3, 3), (5, 9), (7, 1), (10, 14), (14, 17), (16, 16)])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_band_widths_cover_every_kernel(kl, ku, dtype, rng, solve_mode):
    """Every compiled (kl, wu) bucket, with pivoting exercised (no diagonal dominance); the
    reference's relative rule (tests/cross_backend/test_banded_level.py:71-76)."""
    from conatus_amd.runtime.hip import BandedBatch, to_device
    n, r, B = 150, 3, 5
    print(f"case kl={kl} ku={ku} dtype={np.dtype(dtype).name}", flush=True)
    mats = []
    for _ in range(B):
        A = sparse_banded(rng, n, kl, ku, r, dtype, density=0.7)
        A[np.arange(r, n), np.arange(r, n)] += 0.5          # pivots still swap rows
        mats.append(sp.csr_matrix(A))
    bb = BandedBatch(mats, list(range(r)))
    assert bb.solve_mode == solve_mode or (solve_mode == "part" and bb.solve_mode == "wave")
    b = _rhs(rng, (B, n), dtype)
    d = to_device(b)
    bb.solve(d)
    x = d.cpu().numpy()
    _check_against_oracle(mats, list(range(r)), b, x, dtype, well_posed=False)


@pytest.mark.parametrize("kl,ku", [(1, 0), (2, 7), (3, 3), (5, 9), (7, 1), (10, 14), (14, 17), (16, 16)])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_band_widths_conditioned(kl, ku, dtype, rng):
    """Every compiled (kl, wu) bucket on the 'conditioned' family (diagonal + 8): the 1e-13
    backward-error bar."""
    from conatus_amd.runtime.hip import BandedBatch, to_device
    n, r, B = 150, 3, 5
    mats = [conditioned(rng, n, kl, ku, r, shift=8.0 + kl + ku, dtype=complex if dtype == np.complex128 else float)
            for _ in range(B)]
    bb = BandedBatch(mats, list(range(r)))
    b = _rhs(rng, (B, n), dtype)
    d = to_device(b)
    bb.solve(d)
    _check_against_oracle(mats, list(range(r)), b, d.cpu().numpy(), dtype)


def test_single_system_multiple_rhs(rng):
    from conatus_amd.runtime import BorderedBanded
    n = 60
    A, border = bordered_case(rng, n, 3, 2, 2, np.complex128)
    s = BorderedBanded(sp.csr_matrix(A), border)
    assert s.rank == 2 and s.kl <= 3 and s.ku <= 2
    b = _rhs(rng, (n, 3), np.complex128)
    x = s.solve(b)
    assert np.max(abs(A @ x - b)) < 1e-12 * np.max(abs(b)) * n
    x1 = s.solve(b[:, 0])
    assert np.array_equal(x1, x[:, 0])                    # same lane arithmetic for every rhs
    o = OracleBB(sp.csr_matrix(A), border)
    assert abs(s.capacitance_cond - o.cond) <= 1e-8 * o.cond


def test_singular_band_is_refused():
    from conatus_amd.runtime import BorderedBanded
    A = np.eye(6)
    A[3, 3] = 0.0
    with pytest.raises(RuntimeError, match="singular"):
        BorderedBanded(sp.csr_matrix(A), [0])


def test_singular_capacitance_is_refused():
    from conatus_amd.runtime import BorderedBanded
    n = 10
    A = np.eye(n) * 3 + np.diag(np.ones(n - 1), 1)
    A[0] = A[1]                                            # border row dependent on a bulk row
    with pytest.raises(RuntimeError, match="singular"):
        BorderedBanded(sp.csr_matrix(A), [0])


def test_batch_matches_individual(rng):
    from conatus_amd.runtime import BorderedBanded, BorderedBandedBatch
    n, mats, bs = 40, [], []
    for _ in range(6):
        mats.append(sp.csr_matrix(almost_banded(rng, n, 2, 2, [0, 1], complex)))
        bs.append(rng.standard_normal(n) + 1j * rng.standard_normal(n))
    X = BorderedBandedBatch(mats, [0, 1]).solve(np.array(bs))
    for A, b_, x_ in zip(mats, bs, X):
        assert np.max(abs(A @ x_ - b_)) < 1e-11
        assert np.array_equal(BorderedBanded(A, [0, 1]).solve(b_), x_)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_solve_mapped_gather_scatter(dtype, rng):
    torch = _torch()
    from conatus_amd.runtime.hip import BandedBatch
    n, batch, r = 120, 5, 2
    cplx = dtype == np.complex128
    mats = [conditioned(rng, n, 4, 6, r, dtype=complex if cplx else float) for _ in range(batch)]
    bb = BandedBatch(mats, list(range(r)))
    N = batch * n + 17
    rhs = _rhs(rng, N, dtype)
    rows = rng.permutation(N)[: batch * n].reshape(batch, n)
    cols = rng.permutation(N)[: batch * n].reshape(batch, n)
    x = torch.full((N,), 7.0, dtype=torch.complex128 if cplx else torch.float64, device="cuda")
    bb.solve_mapped(torch.from_numpy(rhs.astype(dtype)).cuda(), torch.from_numpy(rows).cuda(), x,
                    torch.from_numpy(cols).cuda())
    xh = x.cpu().numpy()
    for k, A in enumerate(mats):
        xk = xh[cols[k]]
        assert np.max(abs(A @ xk - rhs[rows[k]])) < 1e-10
        assert backward_error(A, xk, rhs[rows[k]]) <= BE_TOL
    untouched = np.setdiff1d(np.arange(N), cols.ravel())
    assert np.all(xh[untouched] == 7.0)
    # contiguous maps (None) equal the plain in-place solve bit for bit
    b = _rhs(rng, (batch, n), dtype)
    d1 = torch.from_numpy(b).cuda()
    d2 = torch.empty_like(d1)
    bb.solve_mapped(d1.reshape(-1), None, d2.reshape(-1), None)
    bb.solve(d1)
    assert torch.equal(d1, d2)


def test_solve_mapped_refuses_overlap_with_different_maps(rng):
    torch = _torch()
    from conatus_amd.runtime.hip import BandedBatch
    n, batch = 30, 2
    mats = [conditioned(rng, n, 2, 2, 1, dtype=float) for _ in range(batch)]
    bb = BandedBatch(mats, [0])
    buf = torch.zeros(batch * n, dtype=torch.float64, device="cuda")
    rows = torch.arange(batch * n, device="cuda").reshape(batch, n)
    cols = torch.flip(rows, dims=[1]).contiguous()
    with pytest.raises(ValueError, match="overlap"):
        bb.solve_mapped(buf, rows, buf, cols)
    with pytest.raises(ValueError, match="out of range"):
        bb.solve_mapped(buf, rows + 1, torch.zeros_like(buf), rows)


def test_solve_split_complex_of_real_data(rng):
    torch = _torch()
    from conatus_amd.runtime.hip import BandedBatch
    n, batch, r = 64, 4, 2
    mats = [conditioned(rng, n, 3, 5, r, dtype=complex) for _ in range(batch)]
    bb = BandedBatch(mats, list(range(r)))
    N = 2 * batch * n + 5
    re = rng.standard_normal(N)
    perm = rng.permutation(N)
    rows_c = perm[: batch * n].reshape(batch, n)
    rows_s = perm[batch * n: 2 * batch * n].reshape(batch, n)
    perm2 = rng.permutation(N)
    cols_c = perm2[: batch * n].reshape(batch, n)
    cols_s = perm2[batch * n: 2 * batch * n].reshape(batch, n)
    x = torch.zeros(N, dtype=torch.float64, device="cuda")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()   # noqa: E731
    bb.solve_split(T(re), T(rows_c), T(rows_s), x, T(cols_c), T(cols_s))
    xh = x.cpu().numpy()
    for k, A in enumerate(mats):
        b = re[rows_c[k]] + 1j * re[rows_s[k]]
        xk = xh[cols_c[k]] + 1j * xh[cols_s[k]]
        assert backward_error(A, xk, b) <= BE_TOL


def test_solves_are_bitwise_reproducible(rng):
    torch = _torch()
    from conatus_amd.runtime.hip import BandedBatch, to_device
    mats = [conditioned(rng, 200, 5, 7, 3, dtype=complex) for _ in range(9)]
    bb = BandedBatch(mats, [0, 1, 2])
    b = _rhs(rng, (9, 200), np.complex128)
    d1, d2 = to_device(b), to_device(b)
    bb.solve(d1)
    bb.solve(d2)
    assert torch.equal(d1, d2)
    bb.refactor()
    d3 = to_device(b)
    bb.solve(d3)
    assert torch.equal(d1, d3)


def test_microbench_systems_full_batch(rng):
    """The microbench batch at full size: B = 4096 complex systems, N = 256, kl = ku = 4, two
    boundary rows of +-1, 4 right-hand sides; backward error of every system <= 1e-13 and oracle
    agreement on a sample."""
    torch = _torch()
    from conatus_amd.runtime.banded import BatchHandle
    B, N, nrhs = 4096, 256, 4
    indptr, indices, datas = microbench_csr(rng, B, N)
    h = BatchHandle([(N, indptr, indices, d) for d in datas], [0, 1], np.complex128)
    p = h.props()
    assert (p.kl, p.ku, p.rank) == (4, 4, 2)
    b = _rhs(rng, (nrhs, B, N), np.complex128)
    d = torch.from_numpy(b).cuda()
    h.solve_ptr(d.data_ptr(), nrhs, B * N, False, torch.cuda.current_stream().cuda_stream)
    x = d.cpu().numpy()
    # every system, vectorised residual: A x - b with the CSR pieces
    rowidx = np.repeat(np.arange(N), np.diff(indptr))
    vals = np.stack(datas)                                  # (B, nnz)
    for k in range(nrhs):
        Ax = np.zeros((B, N), complex)
        np.add.at(Ax.T, rowidx, (vals * x[k][:, indices]).T)
        res = np.abs(Ax - b[k]).max(axis=1)
        nA = np.zeros((B, N))
        np.add.at(nA.T, rowidx, np.abs(vals).T)
        be = res / (nA.max(axis=1) * np.abs(x[k]).max(axis=1) + np.abs(b[k]).max(axis=1))
        assert be.max() <= BE_TOL, be.max()
    for m in rng.choice(B, size=8, replace=False):
        A = csr_of(indptr, indices, datas[m], N)
        xo = OracleBB(A, [0, 1]).solve(b[:, m, :])
        assert np.max(abs(x[:, m, :] - xo)) <= 1e-11 * np.max(abs(xo))


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("r", [0, 2, 4])
def test_transposed_solve_matches_oracle(dtype, r, rng):
    """A_k^T x = b_k with the same factors (banded.hpp:346-372): backward error of A^T and
    agreement with the oracle's transposed solve."""
    from conatus_amd.runtime import BorderedBandedBatch
    n, kl, ku = int(rng.integers(30, 90)), int(rng.integers(1, 5)), int(rng.integers(0, 6))
    print(f"case n={n} kl={kl} ku={ku} r={r}", flush=True)
    mats, border = _batch(rng, 5, n, kl, ku, r, dtype)
    bb = BorderedBandedBatch(mats, border)
    b = _rhs(rng, (5, n), dtype)
    x = bb.solve(b, transpose=True)
    for k, A in enumerate(mats):
        At = sp.csr_matrix(A.T)
        assert backward_error(At, x[k], b[k]) <= BE_TOL
        o = OracleBB(A, border, dtype)
        xo = o.solve_transpose(b[k])
        assert np.max(abs(x[k] - xo)) <= 1e-12 * max(1.0, o.cond) * np.max(abs(xo)) * n


def test_transposed_device_api(rng):
    torch = _torch()
    from conatus_amd.runtime.hip import BandedBatch, to_device
    mats = [conditioned(rng, 120, 4, 6, 2, dtype=complex) for _ in range(6)]
    bb = BandedBatch(mats, [0, 1])
    bb.prepare_transpose()
    b = _rhs(rng, (6, 120), np.complex128)
    d = to_device(b)
    bb.solve_transpose(d)
    x = d.cpu().numpy()
    for k, A in enumerate(mats):
        assert backward_error(sp.csr_matrix(A.T), x[k], b[k]) <= BE_TOL
    # forward after transposed: unaffected
    d2 = to_device(b)
    bb.solve(d2)
    for k, A in enumerate(mats):
        assert backward_error(A, d2.cpu().numpy()[k], b[k]) <= BE_TOL


def test_core_shim_signatures(rng):
    """The _core stand-in: the nanobind signatures of core/bindings/module.cpp:131-206."""
    from conatus_amd import core_shim as C
    from conatus_amd.runtime.banded import csr_args
    n = 50
    mats, border = _batch(rng, 4, n, 3, 2, 2, np.complex128)
    bb = C.BorderedBandedBatchZ([csr_args(A, np.complex128) for A in mats], border, False)
    assert len(bb) == 4 and bb.kl <= 3 and bb.ku <= 2
    b = _rhs(rng, (4, n), np.complex128)
    x = bb.solve(b, False, 4)
    for k, A in enumerate(mats):
        assert backward_error(A, x[k], b[k]) <= BE_TOL
    s = C.BorderedBandedZ(*csr_args(mats[0], np.complex128), border)
    xs = s.solve(b[:1])
    assert np.array_equal(xs[0], x[0]) and s.rank == 2
    with pytest.raises(ValueError):
        bb.solve(b[:3])


def test_auto_mode_selection(rng, monkeypatch, solve_mode):
    """Without CNHIP_MODE: partitioned wave mode for a few systems long enough to split, lane mode
    for >= 1024 narrow-band ones."""
    if solve_mode != "lane" or os.environ.get("CNHIP_SPX"):
        pytest.skip("mode selection is checked once")
    monkeypatch.delenv("CNHIP_MODE")
    from conatus_amd.runtime.hip import BandedBatch
    few = BandedBatch([conditioned(rng, 64, 2, 3, 2, dtype=complex) for _ in range(3)], [0, 1])
    assert few.solve_mode == "part"
    A = conditioned(rng, 32, 2, 2, 2, dtype=complex)
    many = BandedBatch([A] * 1024, [0, 1])
    assert many.solve_mode == "lane"
    assert "lane per (system, rhs)" in many.explain() and "partitioned over" in few.explain()


def test_wave_block_length_requests(rng, solve_mode):
    """wave8 / wave32 (cn_bb_request_mode 4 / 5, the autotuner's candidates): wave mode with staged
    blocks of 8 or of 32 steps; the block length only decides when rows are staged, so the
    solutions are bit for bit the same, and within the oracle bar."""
    if solve_mode != "wave":
        pytest.skip("checked once")
    from conatus_amd.runtime.hip import BandedBatch, to_device
    n, r = 150, 7
    mats = [conditioned(rng, n, 12, 18, r, shift=38.0, dtype=complex) for _ in range(6)]
    b = _rhs(rng, (6, n), np.complex128)
    out = []
    for mode in ("wave8", "wave32"):
        bb = BandedBatch(mats, list(range(r)), mode=mode)
        assert bb.solve_mode == "wave"
        d = to_device(b)
        bb.solve(d)
        out.append(d.cpu().numpy())
    assert np.array_equal(out[0], out[1])
    _check_against_oracle(mats, list(range(r)), b, out[0], np.complex128)


@pytest.mark.parametrize("kl,ku,r", [(12, 18, 7), (4, 4, 2), (16, 16, 3)])
def test_many_wide_band_systems(kl, ku, r, rng, solve_mode):
    """Enough systems that the wave-mode sweeps take their short staged blocks (more than 512
    wavefronts per launch: solve and border completion) at the band of the RB 2D and channel mode
    systems; the 1e-13 backward-error bar on a sample checked against the oracle."""
    from conatus_amd.runtime.hip import BandedBatch, to_device
    if solve_mode != "wave":
        pytest.skip("the wave-mode block length is what this case exercises")
    n, B = 120, 640
    print(f"case kl={kl} ku={ku} r={r} B={B}", flush=True)
    base = [conditioned(rng, n, kl, ku, r, shift=8.0 + kl + ku, dtype=complex) for _ in range(8)]
    mats = [base[k % 8] for k in range(B)]
    bb = BandedBatch(mats, list(range(r)))
    assert bb.solve_mode == "wave"
    b = _rhs(rng, (B, n), np.complex128)
    d = to_device(b)
    bb.solve(d)
    x = d.cpu().numpy()
    sample = sorted(set(rng.choice(B, size=12, replace=False).tolist()) | {0, B - 1})
    _check_against_oracle([mats[k] for k in sample], list(range(r)), b[sample], x[sample], np.complex128)
