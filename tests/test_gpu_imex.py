"""GPU parity of the IMEX stage (fused SBDF/RK right-hand side + batched solve) against the CPU
oracle (oracle/imex.py, restating core/include/conatus/imex.hpp).

Bar: the assembled right-hand sides and CSR products equal the oracle's bit for bit (same
rounded operations, no contraction); states after 100 steps within 1e-11 relative (the solves
differ from LAPACK's only by rounding); graph replays equal eager steps bit for bit; run-to-run
determinism (as tests/cross_backend/test_gpu_vs_cpu.py:151-178)."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.imex import ImexIntegrator, sbdf_coefficients
from tests.imex_cases import heat_system

pytestmark = pytest.mark.gpu

SCHEMES = ["SBDF1", "SBDF2", "SBDF3", "SBDF4", "RK111", "RK222", "RK443", "RKSMR", "RKGFY",
           "CNAB1", "CNAB2", "MCNAB2", "CNLF2"]


def rel(a, b):
    return np.max(abs(a - b)) / max(np.max(abs(b)), 1e-300)


def _problem(rng, N=32, batch=6, dtype=np.float64):
    kappas = np.linspace(0.0, 5.0, batch)
    systems = [heat_system(N, k) for k in kappas]
    n = N + 2
    cplx = np.dtype(dtype) == np.complex128
    def r(shape):
        a = rng.standard_normal(shape)
        return a + 1j * rng.standard_normal(shape) if cplx else a
    g = 0.3 * r((batch, n))
    h = 0.1 * r((batch, n))
    g[:, N:] = 0.0
    h[:, N:] = 0.0
    x0 = np.zeros((batch, n), dtype=dtype)
    x0[:, :N] = r((batch, N)) * np.exp(-0.3 * np.arange(N))
    Ms = [s[0].astype(dtype) for s in systems]
    Ls = [s[1].astype(dtype) for s in systems]
    return Ms, Ls, systems[0][2], g, h, x0


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_batched_imex_matches_oracle(scheme, dtype, rng):
    import torch
    from conatus_amd.runtime.imex import Imex
    dt, steps = 2e-3, 100
    Ms, Ls, border, g, h, x0 = _problem(rng, dtype=dtype)
    print(f"case {scheme} {np.dtype(dtype).name}", flush=True)
    gpu = Imex(Ms, Ls, border, scheme, dt, g=g, h=h)
    d = torch.from_numpy(x0.copy()).cuda()
    t = gpu.run(d, 0.0, steps)
    x_gpu = d.cpu().numpy()
    assert abs(t - steps * dt) < 1e-12
    for k in range(len(Ms)):
        gk, hk = g[k].copy(), h[k].copy()
        cpu = ImexIntegrator(Ms[k], Ls[k], border, scheme, dt, F=lambda x, t, gk=gk, hk=hk: gk * x + hk,
                             dtype=dtype)
        x_cpu, _ = cpu.run(x0[k], 0.0, steps)
        assert rel(x_gpu[k], x_cpu) < 1e-11, (scheme, k, rel(x_gpu[k], x_cpu))
    if scheme == "RK443":
        assert gpu.n_factorizations == 1
    if scheme == "SBDF4":
        assert gpu.n_factorizations == 4
    d2 = torch.from_numpy(x0.copy()).cuda()
    Imex(Ms, Ls, border, scheme, dt, g=g, h=h).run(d2, 0.0, steps)
    assert torch.equal(d2, d)                                 # run-to-run determinism


@pytest.mark.parametrize("scheme", ["SBDF2", "SBDF3", "RK443"])
def test_graph_replay_equals_eager_steps(scheme, rng):
    import torch
    from conatus_amd.runtime.imex import Imex
    Ms, Ls, border, g, h, x0 = _problem(rng)
    a = Imex(Ms, Ls, border, scheme, 1e-3, g=g, h=h)
    b = Imex(Ms, Ls, border, scheme, 1e-3, g=g, h=h)
    da = torch.from_numpy(x0.copy()).cuda()
    db = torch.from_numpy(x0.copy()).cuda()
    ta = a.run(da, 0.0, 23)
    tb = b.run_graph(db, 0.0, 10)
    tb = b.run_graph(db, tb, 13)                               # continuation across calls
    torch.cuda.synchronize()
    assert ta == tb
    assert torch.equal(da, db)


def test_sbdf_stage_rhs_is_bitwise_the_oracle(rng):
    """One fused stage against numpy: M x by scipy, rhs = sum_j (c_j F_j - a_j/dt MX_j)."""
    import torch
    from conatus_amd import _lib
    from conatus_amd.runtime.imex import DeviceCsr
    for dtype in (np.float64, np.complex128):
        Ms, Ls, border, g, h, x0 = _problem(rng, dtype=dtype)
        B, n = x0.shape
        cplx = dtype == np.complex128
        Md = DeviceCsr(Ms, dtype)
        q, dt = 3, 7e-4
        a, c = sbdf_coefficients(q)
        hist_mx = [rng.standard_normal((B, n)) + (1j * rng.standard_normal((B, n)) if cplx else 0) for _ in range(2)]
        hist_f = [rng.standard_normal((B, n)) + (1j * rng.standard_normal((B, n)) if cplx else 0) for _ in range(2)]
        T = lambda a_: torch.from_numpy(np.ascontiguousarray(a_, dtype=dtype)).cuda()   # noqa: E731
        x, gd, hd = T(x0), T(g), T(h)
        mx0, f0, rhs = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
        hm, hf = [T(v) for v in hist_mx], [T(v) for v in hist_f]
        P = ctypes.c_void_p * 4
        v = ctypes.c_void_p
        st = _lib.lib().cn_sbdf_stage(Md._h, v(x.data_ptr()), v(gd.data_ptr()), v(hd.data_ptr()), None,
                                      v(mx0.data_ptr()), v(f0.data_ptr()),
                                      P(hm[0].data_ptr(), hm[1].data_ptr(), 0, 0),
                                      P(hf[0].data_ptr(), hf[1].data_ptr(), 0, 0), q,
                                      (ctypes.c_double * 5)(0, c[1], c[2], c[3], 0),
                                      (ctypes.c_double * 5)(0, a[1] / dt, a[2] / dt, a[3] / dt, 0),
                                      v(rhs.data_ptr()), v(torch.cuda.current_stream().cuda_stream))
        _lib.check(st, "test")
        torch.cuda.synchronize()
        MX = np.stack([Ms[k] @ x0[k] for k in range(B)])      # scipy: in-order, separately rounded
        # F = g x + h with the separately rounded complex product of std::complex (numpy's own
        # complex multiply may fuse)
        if cplx:
            F = (g.real * x0.real - g.imag * x0.imag) + 1j * (g.real * x0.imag + g.imag * x0.real)
            F = (F.real + h.real) + 1j * (F.imag + h.imag)
        else:
            F = g * x0 + h
        levels_f = [F] + hist_f
        levels_mx = [MX] + hist_mx
        R = np.zeros((B, n), dtype)
        for j in range(1, q + 1):              # real scalars times values: no fusion possible
            R = R + (c[j] * levels_f[j - 1] - (a[j] / dt) * levels_mx[j - 1])
        assert np.array_equal(mx0.cpu().numpy(), MX)
        assert np.array_equal(f0.cpu().numpy(), F)
        assert np.array_equal(rhs.cpu().numpy(), R)


def test_continuation_rule(rng):
    import torch
    from conatus_amd.errors import OptionError
    from conatus_amd.runtime.imex import Imex
    Ms, Ls, border, g, h, x0 = _problem(rng)
    im = Imex(Ms, Ls, border, "SBDF2", 1e-3, g=g, h=h)
    d = torch.from_numpy(x0.copy()).cuda()
    t = im.run(d, 0.0, 3)
    im.run(d, t, 2)
    with pytest.raises(OptionError, match="history"):
        im.run(torch.from_numpy(x0.copy()).cuda(), 0.0, 1)
    im.reset()
    im.run(torch.from_numpy(x0.copy()).cuda(), 0.0, 1)
