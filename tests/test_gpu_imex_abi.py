"""GPU tests of the IMEX integrator handle of the C ABI (cn_imex_*, the reference's cuda.Imex for
callers that bind libcnhip directly), called through ctypes with torch device memory.

Bar: states after 100 steps within 1e-11 relative of the CPU oracle (oracle/imex.py, restating
core/include/conatus/imex.hpp) and bit for bit those of conatus_amd.runtime.imex.Imex (the same
library calls in the same order); the reference's continuation rule (a foreign (x, t0) refused
while a history is held, reset starts over) and its clock (t0 + one dt per step)."""
import ctypes

import numpy as np
import pytest

from oracle.imex import ImexIntegrator
from tests.imex_cases import heat_system

pytestmark = pytest.mark.gpu


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
    return [s[0].astype(dtype) for s in systems], [s[1].astype(dtype) for s in systems], systems[0][2], g, h, x0


class _AbiImex:
    """ctypes owner of one cn_imex handle (what a C caller writes)."""

    def __init__(self, Ms, Ls, border, scheme, dt, dtype):
        import torch
        from conatus_amd import _lib
        from conatus_amd.runtime.banded import csr_args
        self.L = _lib.lib()
        self._lib = _lib
        B, n = len(Ms), Ms[0].shape[0]
        self.keep = []

        def arrays(mats):
            cs = [csr_args(A.tocsr(), dtype) for A in mats]
            self.keep.append(cs)
            P = ctypes.c_void_p * B
            return (P(*[c[1].ctypes.data for c in cs]), P(*[c[2].ctypes.data for c in cs]),
                    P(*[np.ascontiguousarray(c[3]).ctypes.data for c in cs]),
                    (ctypes.c_int64 * B)(*[int(c[2].size) for c in cs]))
        m, l = arrays(Ms), arrays(Ls)
        bord = np.asarray(border, np.int32)
        self.h = ctypes.c_void_p()
        code = _lib.CN_C128 if np.dtype(dtype) == np.complex128 else _lib.CN_F64
        _lib.check(self.L.cn_imex_create(ctypes.byref(self.h), code, B, n, *m, *l,
                                         bord.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(bord.size),
                                         scheme.encode(), float(dt), 1), "cn_imex_create")
        self.stream = torch.cuda.current_stream().cuda_stream

    def set_pointwise(self, g, h):
        self._lib.check(self.L.cn_imex_set_pointwise(self.h, g.data_ptr(), h.data_ptr(), self.stream), "set_pointwise")

    def run(self, d, t0, steps):
        t = ctypes.c_double()
        st = self.L.cn_imex_run(self.h, d.data_ptr(), float(t0), int(steps), ctypes.byref(t), self.stream)
        return st, t.value

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        self._lib.check(self.L.cn_imex_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "info")
        return a.value, b.value, c.value

    def __del__(self):
        if getattr(self, "h", None):
            self.L.cn_imex_destroy(self.h)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("scheme", ["SBDF1", "SBDF2", "SBDF3", "SBDF4", "RK111", "RK222", "RK443"])
def test_abi_integrator_matches_oracle_and_python_imex(scheme, dtype, rng):
    import torch
    from conatus_amd import _lib
    from conatus_amd.runtime.imex import Imex
    dt, steps = 2e-3, 100
    Ms, Ls, border, g, h, x0 = _problem(rng, dtype=dtype)
    print(f"case cn_imex {scheme} {np.dtype(dtype).name}", flush=True)
    abi = _AbiImex(Ms, Ls, border, scheme, dt, dtype)
    abi.set_pointwise(torch.from_numpy(g.astype(dtype)).cuda(), torch.from_numpy(h.astype(dtype)).cuda())
    d = torch.from_numpy(x0.copy()).cuda()
    st, t = abi.run(d, 0.0, 60)
    _lib.check(st, "cn_imex_run")
    st, t = abi.run(d, t, steps - 60)                         # continuation of the same trajectory
    _lib.check(st, "cn_imex_run")
    torch.cuda.synchronize()
    x_abi = d.cpu().numpy()
    assert t == pytest.approx(steps * dt, abs=1e-12)
    py = Imex(Ms, Ls, border, scheme, dt, g=g, h=h)
    dp = torch.from_numpy(x0.copy()).cuda()
    py.run(dp, 0.0, steps)
    assert np.array_equal(x_abi, dp.cpu().numpy())
    for k in range(len(Ms)):
        gk, hk = g[k].copy(), h[k].copy()
        o = ImexIntegrator(Ms[k], Ls[k], border, scheme, dt, F=lambda x, tt, gk=gk, hk=hk: gk * x + hk, dtype=dtype)
        xo, _ = o.run(x0[k].copy(), 0.0, steps)
        assert rel(x_abi[k], xo) < 1e-11, (k, rel(x_abi[k], xo))
    hist, nfac, it = abi.info()
    assert it == steps and nfac >= 1 and hist == (int(scheme[4]) if scheme.startswith("SBDF") else 0)


def test_abi_continuation_rule_and_reset(rng):
    import torch
    from conatus_amd import _lib
    Ms, Ls, border, g, h, x0 = _problem(rng)
    abi = _AbiImex(Ms, Ls, border, "SBDF2", 1e-3, np.float64)
    d = torch.from_numpy(x0.copy()).cuda()
    st, t = abi.run(d, 0.0, 5)
    assert st == _lib.CN_OK
    st, _ = abi.run(torch.from_numpy(x0.copy()).cuda(), 0.0, 1)     # a foreign state with a history held
    assert st == _lib.CN_EINVAL and "history" in _lib.last_error()
    st, _ = abi.run(d, t + 1.0, 1)                                   # the right state at the wrong time
    assert st == _lib.CN_EINVAL
    _lib.check(abi.L.cn_imex_reset(abi.h), "reset")
    st, t2 = abi.run(torch.from_numpy(x0.copy()).cuda(), 0.0, 5)
    assert st == _lib.CN_OK and t2 == t
    bad = ctypes.c_void_p()
    z = ctypes.c_void_p * 1
    n64 = (ctypes.c_int64 * 1)(0)
    st = abi.L.cn_imex_create(ctypes.byref(bad), _lib.CN_F64, 1, 4, z(0), z(0), z(0), n64, z(0), z(0), z(0), n64,
                              None, 0, b"SBDF9", 1e-3, 1)
    assert st == _lib.CN_EINVAL and "scheme" in _lib.last_error()
