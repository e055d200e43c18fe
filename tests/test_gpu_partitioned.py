"""Partitioned wave mode: the forward and backward sweeps of one system split over 2, 4 or 8
wavefronts of a workgroup, each sweeping its chunk with the coupling rows from its neighbour set to
zero, the chunks then coupled through spikes computed at factor time (the chunk's response to unit
coupling values).  Checked against the oracle (the reference's gbtrs-order solve), against the
one-wavefront solve of the same factors, and for batch independence (a system solves bit for bit
alike in any batch of this mode)."""
import numpy as np
import pytest
import scipy.sparse as sp

from tests.banded_cases import backward_error, conditioned, sparse_banded

pytestmark = pytest.mark.gpu

BE_TOL = 1e-13


def _rhs(rng, shape, dtype):
    b = rng.standard_normal(shape)
    if np.dtype(dtype) == np.complex128:
        b = b + 1j * rng.standard_normal(shape)
    return b


def _mats(rng, B, n, kl, ku, r, dtype, pivoting):
    if pivoting:          # no diagonal dominance: partial pivoting swaps rows
        out = []
        for _ in range(B):
            A = sparse_banded(rng, n, kl, ku, r, dtype, density=0.7)
            A[np.arange(r, n), np.arange(r, n)] += 0.5
            out.append(sp.csr_matrix(A))
        return out
    return [conditioned(rng, n, kl, ku, r, shift=8.0 + kl + ku, dtype=complex if dtype == np.complex128 else float)
            for _ in range(B)]


def _solve(mats, r, b, monkeypatch, mode, parts=None):
    from conatus_amd.runtime.hip import BandedBatch, to_device
    monkeypatch.setenv("CNHIP_MODE", mode)
    if parts:
        monkeypatch.setenv("CNHIP_PARTS", str(parts))
    else:
        monkeypatch.delenv("CNHIP_PARTS", raising=False)
    bb = BandedBatch(mats, list(range(r)))
    if b.ndim == 3:       # (nrhs, batch, n)
        import torch
        d = torch.from_numpy(np.ascontiguousarray(b)).cuda()
        bb.solve_many(d)
    else:
        d = to_device(b)
        bb.solve(d)
    return bb, d.cpu().numpy()


@pytest.mark.parametrize("parts", [2, 4, 8])
@pytest.mark.parametrize("kl,ku,r", [(12, 18, 7), (4, 4, 2), (2, 3, 0), (1, 0, 1), (8, 20, 3), (16, 16, 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_parts_match_oracle(parts, kl, ku, r, dtype, rng, monkeypatch):
    """n = 263 (the RB 2D 128x64 mode systems' size; uneven chunks), three right-hand sides per
    system: every solution within the 1e-13 backward-error bar."""
    from oracle.banded import BorderedBanded as OracleBB
    n, B = 263, 5
    print(f"case parts={parts} kl={kl} ku={ku} r={r} dtype={np.dtype(dtype).name}", flush=True)
    mats = _mats(rng, B, n, kl, ku, r, dtype, pivoting=False)
    b = _rhs(rng, (3, B, n), dtype)
    bb, x = _solve(mats, r, b, monkeypatch, "part", parts)
    assert bb.solve_mode == "part"
    for k, A in enumerate(mats):
        o = OracleBB(A, list(range(r)), dtype)
        for q in range(3):
            assert backward_error(A, x[q, k], b[q, k]) <= BE_TOL, (k, q)
            xo = o.solve(b[q, k])
            assert np.max(abs(x[q, k] - xo)) <= 1e-12 * max(1.0, o.cond) * np.max(abs(xo)) * n


@pytest.mark.parametrize("parts", [2, 4, 8])
@pytest.mark.parametrize("kl,ku,r", [(12, 18, 7), (3, 3, 2), (7, 1, 0)])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_parts_with_row_interchanges_match_wave_mode(parts, kl, ku, r, dtype, rng, monkeypatch):
    """Bands without diagonal dominance (pivots swap rows, also across chunk boundaries), the
    reference's cross-backend rule against the one-wavefront solve of the same factors (the
    reference's gbtrs order): where that solve's backward error is within 1e-12 (these random bands
    are at times nearly singular, and then no order of the sweeps has a bar to meet), the
    partitioned solve's is within 10x of it or 1e-13."""
    n, B = 263, 4
    print(f"case parts={parts} kl={kl} ku={ku} r={r} dtype={np.dtype(dtype).name}", flush=True)
    mats = _mats(rng, B, n, kl, ku, r, dtype, pivoting=True)
    b = _rhs(rng, (B, n), dtype)
    bp, xp = _solve(mats, r, b, monkeypatch, "part", parts)
    bw, xw = _solve(mats, r, b, monkeypatch, "wave")
    assert bp.solve_mode == "part" and bw.solve_mode == "wave"
    checked = 0
    for k, A in enumerate(mats):
        be_w = backward_error(A, xw[k], b[k])
        if be_w > 1e-12:
            continue
        be_p = backward_error(A, xp[k], b[k])
        assert be_p <= max(BE_TOL, 10 * be_w), (k, be_p, be_w)
        checked += 1
    print(f"{checked} of {B} systems within the bar of the one-wavefront solve", flush=True)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_parts_bitwise_independent_of_batch(dtype, rng, monkeypatch):
    """Three systems solved alone and placed among six others: identical bits."""
    n, kl, ku, r = 263, 12, 18, 7
    mats = _mats(rng, 9, n, kl, ku, r, dtype, pivoting=True)
    b = _rhs(rng, (9, n), dtype)
    _, x_all = _solve(mats, r, b, monkeypatch, "part", 4)
    sel = [7, 2, 5]
    _, x_sel = _solve([mats[i] for i in sel], r, b[sel], monkeypatch, "part", 4)
    assert np.array_equal(x_all[sel], x_sel)


def test_automatic_selection_by_batch_size(rng, monkeypatch):
    """Without CNHIP_MODE: the RB 2D 128x64 shape (n = 263, kl = 12, ku = 18, r = 7) is
    partitioned for 64 systems (eight wavefronts each: chunks of 32-33 rows against a coupling
    width of 30), one wavefront each for 300; a system too short to split stays in wave mode."""
    from conatus_amd.runtime.hip import BandedBatch
    monkeypatch.delenv("CNHIP_MODE", raising=False)
    monkeypatch.delenv("CNHIP_PARTS", raising=False)
    A = conditioned(rng, 263, 12, 18, 7, shift=38.0, dtype=complex)
    assert BandedBatch([A] * 64, list(range(7))).solve_mode == "part"
    assert BandedBatch([A] * 300, list(range(7))).solve_mode == "wave"
    short = conditioned(rng, 60, 12, 18, 7, shift=38.0, dtype=complex)
    assert BandedBatch([short] * 4, list(range(7))).solve_mode == "wave"
    assert "partitioned" in BandedBatch([A] * 2, list(range(7))).explain()


def test_copy_systems_carries_the_spikes(rng, monkeypatch):
    """Factored systems moved between two partitioned handles (cn_bb_copy_systems) solve exactly as
    in the handle that factored them."""
    import torch

    from conatus_amd.runtime.hip import BandedBatch, to_device
    monkeypatch.setenv("CNHIP_MODE", "part")
    n, r = 263, 7
    src_m = _mats(rng, 3, n, 12, 18, r, np.complex128, pivoting=True)
    dst_m = _mats(rng, 5, n, 12, 18, r, np.complex128, pivoting=True)
    src = BandedBatch(src_m, list(range(r)))
    dst = BandedBatch(dst_m, list(range(r)))
    assert src.solve_mode == dst.solve_mode == "part"
    dst._h.copy_systems_from(src._h, 0, 1, 3, torch.cuda.current_stream().cuda_stream)
    b = _rhs(rng, (5, n), np.complex128)
    d = to_device(b)
    dst.solve(d)
    e = to_device(b[1:4])
    src.solve(e)
    assert np.array_equal(d.cpu().numpy()[1:4], e.cpu().numpy())
