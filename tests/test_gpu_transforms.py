"""GPU parity of the batched transforms (conatus_amd.transforms, libcnhip cn_tr_rows) against the
CPU oracle's FFT / DCT transforms (oracle/spectral.py, itself pinned by closed forms).

Bar: relative to the largest input entry, forward transforms within 1e-13 and backward ones within
1e-12 (dense FP64 sums of up to a few hundred terms against FFT rounding); a coefficient -> grid ->
coefficient round trip is the identity to 1e-13; the reference's error types for wrong shapes and
data types.  Sizes include the dealiased grids of this path (Chebyshev 97 -> 145, Fourier
128 -> 192) and a batch of 4096 rows."""
import numpy as np
import pytest

from oracle import spectral as S

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel(a, b):
    return float(np.max(abs(a - b)) / max(np.max(abs(b)), 1e-300))


@pytest.mark.parametrize("N,dealias,rows", [(16, 1.5, 7), (97, 1.5, 33), (64, 1.0, 5), (128, 1.5, 4096)])
@pytest.mark.parametrize("cplx", [False, True])
def test_chebyshev_rows_match_the_oracle(N, dealias, rows, cplx, rng):
    from conatus_amd import transforms as TR
    b = TR.Chebyshev(N, (-1.0, 1.0), dealias)
    M = TR.grid_size(b)
    print(f"case chebyshev N={N} M={M} rows={rows} complex={cplx}", flush=True)
    f = rng.standard_normal((rows, M)) + (1j * rng.standard_normal((rows, M)) if cplx else 0)
    c = TR.forward(_dev(f), b).cpu().numpy()
    assert c.shape == (rows, N)
    assert _rel(c, S.cheb_forward(f, N, axis=1)) < 1e-13
    g = TR.backward(_dev(c), b).cpu().numpy()
    assert _rel(g, S.cheb_backward(c, M, axis=1)) < 1e-12
    c2 = TR.forward(_dev(g), b).cpu().numpy()
    assert _rel(c2, c) < 1e-13                                 # round trip on coefficient data


@pytest.mark.parametrize("N,dealias,rows", [(8, 1.5, 3), (64, 1.5, 40), (128, 1.5, 4096), (32, 1.0, 9)])
def test_fourier_rows_match_the_oracle(N, dealias, rows, rng):
    from conatus_amd import transforms as TR
    rb = TR.RealFourier(N, (0.0, 2.0), dealias)
    cb = TR.ComplexFourier(N, (0.0, 2.0), dealias)
    M = TR.grid_size(rb)
    print(f"case fourier N={N} M={M} rows={rows}", flush=True)
    f = rng.standard_normal((rows, M))
    c = TR.forward(_dev(f), rb).cpu().numpy()
    assert _rel(c, S.rf_forward(f, N, axis=1)) < 1e-13
    assert np.all(c[:, 1] == 0.0)                             # msin_0
    g = TR.backward(_dev(c), rb).cpu().numpy()
    assert _rel(g, S.rf_backward(c, M, axis=1)) < 1e-12
    assert _rel(TR.forward(_dev(g), rb).cpu().numpy(), c) < 1e-13
    fz = f + 1j * rng.standard_normal((rows, M))
    cz = TR.forward(_dev(fz), cb).cpu().numpy()
    assert _rel(cz, S.cf_forward(fz, N, axis=1)) < 1e-13
    assert np.all(cz[:, N // 2] == 0)                         # Nyquist slot
    gz = TR.backward(_dev(cz), cb).cpu().numpy()
    assert _rel(gz, S.cf_backward(cz, M, axis=1)) < 1e-12


def test_transform_errors_are_the_references():
    import torch
    from conatus_amd import transforms as TR
    b = TR.Chebyshev(16)
    with pytest.raises(ValueError, match="grid points"):
        TR.forward(torch.zeros((3, 16), dtype=torch.float64, device="cuda"), b)   # needs 24
    with pytest.raises(ValueError, match="coefficients"):
        TR.backward(torch.zeros((3, 24), dtype=torch.float64, device="cuda"), b)
    with pytest.raises(TypeError, match="RealFourier takes real data"):
        TR.forward(torch.zeros((2, 24), dtype=torch.complex128, device="cuda"), TR.RealFourier(16))
    with pytest.raises(TypeError, match="ComplexFourier takes complex"):
        TR.forward(torch.zeros((2, 24), dtype=torch.float64, device="cuda"), TR.ComplexFourier(16))
    with pytest.raises(ValueError, match="even"):
        TR.forward(torch.zeros((2, 22), dtype=torch.float64, device="cuda"), TR.RealFourier(15))
    out = TR.forward(torch.zeros((0, 24), dtype=torch.float64, device="cuda"), b)  # empty batch
    assert tuple(out.shape) == (0, 16)


def test_shim_cuda_transform_bindings(rng):
    """The reference's binding names (cheb_forward(grid, N), realfourier_backward(coef, M), ...)
    through core_shim.cuda on DeviceArrays."""
    from conatus_amd import core_shim as C
    f = rng.standard_normal((5, 24))
    c = C.cuda.cheb_forward(C.cuda.DeviceArrayD.from_host(f), 16)
    assert isinstance(c, C.cuda.DeviceArrayD) and c.shape == (5, 16)
    assert _rel(c.to_host(), S.cheb_forward(f, 16, axis=1)) < 1e-13
    g = C.cuda.realfourier_backward(C.cuda.DeviceArrayD.from_host(S.rf_forward(f, 16, axis=1)), 24)
    assert _rel(g.to_host(), S.rf_backward(S.rf_forward(f, 16, axis=1), 24, axis=1)) < 1e-12
    z = f + 1j * rng.standard_normal((5, 24))
    cz = C.cuda.complexfourier_forward(C.cuda.DeviceArrayZ.from_host(z), 16)
    assert _rel(cz.to_host(), S.cf_forward(z, 16, axis=1)) < 1e-13
    with pytest.raises(TypeError):
        C.cuda.realfourier_forward(C.cuda.DeviceArrayZ.from_host(z), 16)
