"""The ``conatus._core`` stand-in (conatus_amd/core_shim.py) driven the way the reference's Python
layer drives its compiled module: the host ImexIntegrator binding with a Python F(x, t)
callback (python/conatus/runtime/imex.py:39-90 calls ``_core.ImexD/ImexZ(n, M arrays, L arrays,
border, scheme, dt)``, ``set_rhs``, ``run``, ``reset``), and the ``cuda`` submodule's
BandedBatch / Imex / DeviceArray classes with torch pointers for the pointer entry points
(python/conatus/runtime/cuda.py:127-308, core/bindings/module.cpp:321-442).

Bars: integrator states within 1e-11 relative of the CPU oracle (oracle/imex.py restates the
reference's ImexIntegrator, the solves differ by rounding only); solves to a backward error
<= 1e-13 and within 1e-12 of the oracle's; the refusals carry the binding's exception types and
the keywords the reference's tests match ("reset", "singular")."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.banded import BorderedBanded as OracleBB
from oracle.imex import ImexIntegrator as OracleImex
from oracle.spectral import cheb_forward
from tests.banded_cases import backward_error, conditioned
from tests.imex_cases import N, exact, heat_system, make_F

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(abs(a - b)) / max(np.max(abs(b)), 1e-300))


def _reference_style_imex(C, M, L, border, scheme, dt, F, dtype):
    """What python/conatus/runtime/imex.py:39-58 does with ``_core`` (csr arrays, then the
    callback wrapped to return arrays of the integrator's dtype)."""
    from conatus_amd.runtime.banded import csr_args, solve_operand
    cls = C.ImexZ if dtype == np.complex128 else C.ImexD
    n, mp, mi, md = csr_args(M, dtype, "ImexIntegrator (M)")
    _, lp, li, ld = csr_args(L, dtype, "ImexIntegrator (L)")
    s = cls(n, mp, mi, md, lp, li, ld, [int(i) for i in border], scheme, float(dt))
    if F is not None:
        s.set_rhs(lambda x, t: solve_operand(F(np.asarray(x), t), dtype, "ImexIntegrator: F(x, t)"))
    return s


@pytest.mark.parametrize("scheme", ["SBDF1", "SBDF2", "SBDF3", "SBDF4", "RK111", "RK222", "RK443"])
def test_host_imex_with_callback_matches_oracle(scheme):
    from conatus_amd import core_shim as C
    M, L, border = heat_system()
    F = make_F("cos")
    dt, steps = 1e-3, 40
    x0 = np.zeros(N + 2)
    x0[:N] = cheb_forward(exact(0.0), N)
    s = _reference_style_imex(C, M, L, border, scheme, dt, F, np.float64)
    x, t = s.run(x0, 0.0, steps)
    o = OracleImex(M, L, border, scheme, dt, F=F)
    xo, to = o.run(x0.copy(), 0.0, steps)
    print(f"{scheme}: relative difference {rel(x, xo):.2e}", flush=True)
    assert rel(x, xo) < 1e-11
    assert t == to and s.iteration == steps and s.scheme == scheme
    assert s.n_factorizations == o.n_factorizations
    assert s.history_length == o.history_length


def test_host_imex_continuation_and_reset():
    from conatus_amd import core_shim as C
    M, L, border = heat_system()
    rng = np.random.default_rng(3)
    Mz, Lz = M.astype(complex), L.astype(complex)
    s = _reference_style_imex(C, Mz, Lz, border, "SBDF2", 1e-3, None, np.complex128)
    x0 = np.zeros(N + 2, complex)
    x0[:N] = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) * np.exp(-0.4 * np.arange(N))
    x, t = s.run(x0, 0.0, 5)
    x2, t2 = s.run(x, t, 5)                       # continues the returned (x, t)
    o = OracleImex(Mz, Lz, border, "SBDF2", 1e-3)
    xo, _ = o.run(x0.copy(), 0.0, 10)
    assert rel(x2, xo) < 1e-11 and abs(t2 - 0.01) < 1e-15
    with pytest.raises(ValueError, match="reset"):
        s.run(x0, 0.0, 1)
    s.reset()
    x3, _ = s.run(x0, 0.0, 10)
    assert np.array_equal(x3, x2)
    with pytest.raises(ValueError, match="wrong length"):
        s.set_rhs(lambda x, t: np.zeros(3, complex))
        s.reset()
        s.run(x0, 0.0, 1)


def test_cuda_banded_batch_pointer_entry_points(rng):
    import torch
    from conatus_amd import core_shim as C
    from conatus_amd.runtime.banded import csr_args
    n, batch, r = 80, 12, 2
    mats = [conditioned(rng, n, 3, 4, r, dtype=complex) for _ in range(batch)]
    bb = C.cuda.BandedBatchZ([csr_args(A, np.complex128) for A in mats], [0, 1])
    assert (bb.n, bb.batch, bb.rank) == (n, batch, r) and bb.pivot and not bb.level_scheduled
    assert bb.capacitance_cond_max >= 1.0 and bb.factor_ms >= 0.0
    b = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    stream = torch.cuda.current_stream().cuda_stream
    d = torch.from_numpy(b).cuda()
    bb.solve_ptr(d.data_ptr(), batch, n, stream)
    x = d.cpu().numpy()
    for k in range(batch):
        assert backward_error(mats[k], x[k], b[k]) <= 1e-13
        xo = OracleBB(sp.csr_matrix(mats[k]), [0, 1]).solve(b[k])
        assert rel(x[k], xo) < 1e-12
    # transposed, after prepare_transpose (needed before graph capture)
    bb.prepare_transpose()
    d = torch.from_numpy(b).cuda()
    bb.solve_transpose_ptr(d.data_ptr(), batch, n, stream)
    xt = d.cpu().numpy()
    for k in range(batch):
        assert backward_error(sp.csr_matrix(mats[k]).T, xt[k], b[k]) <= 1e-13
    # mapped gather/scatter: b_m[i] = rhs[rows[m, i]], x[cols[m, i]] = x_m[i]
    perm = rng.permutation(batch * n)
    rows = torch.from_numpy(perm.reshape(batch, n).astype(np.int64)).cuda()
    cols = torch.from_numpy(rng.permutation(batch * n).reshape(batch, n).astype(np.int64)).cuda()
    rhs = torch.from_numpy(b.ravel()[np.argsort(perm)].copy()).cuda()    # rhs[rows] = b
    out = torch.zeros(batch * n, dtype=torch.complex128, device="cuda")
    bb.solve_mapped_ptr(rhs.data_ptr(), rows.data_ptr(), out.data_ptr(), cols.data_ptr(), stream)
    got = out.cpu().numpy()[cols.cpu().numpy()]
    assert rel(got, x) < 1e-13
    # the DeviceArray path
    da = C.cuda.DeviceArrayZ.from_host(b)
    bb.solve(da)
    assert np.array_equal(da.to_host(), x)
    with pytest.raises(ValueError):
        bb.solve_ptr(d.data_ptr(), batch + 1, n, stream)
    with pytest.raises(ValueError):
        bb.solve_ptr(0, batch, n, stream)


def test_cuda_banded_batch_singular_is_runtime_error():
    from conatus_amd import core_shim as C
    from conatus_amd.runtime.banded import csr_args
    A = sp.csr_matrix(np.diag([1.0, 0.0, 2.0, 3.0]))
    with pytest.raises(RuntimeError, match="singular"):
        C.cuda.BandedBatchD([csr_args(A, np.float64)], [])


def test_cuda_imex_pointwise_matches_oracle(rng):
    from conatus_amd import core_shim as C
    from conatus_amd.runtime.banded import csr_args
    batch = 4
    systems = [heat_system(N, k) for k in np.linspace(0.0, 3.0, batch)]
    n = N + 2
    g = 0.3 * rng.standard_normal((batch, n))
    h = 0.1 * rng.standard_normal((batch, n))
    g[:, N:] = h[:, N:] = 0.0
    x0 = np.zeros((batch, n))
    x0[:, :N] = rng.standard_normal((batch, N)) * np.exp(-0.3 * np.arange(N))
    im = C.cuda.ImexD([csr_args(s[0], np.float64) for s in systems], [csr_args(s[1], np.float64) for s in systems],
                      systems[0][2], "SBDF3", 1e-3)
    im.set_pointwise_rhs(g, h)
    x = C.cuda.DeviceArrayD.from_host(x0)
    t = im.run(x, 0.0, 30)
    assert im.iteration == 30 and im.history_length == 3 and abs(t - 0.03) < 1e-15
    got = x.to_host()
    for k, (M, L, border) in enumerate(systems):
        o = OracleImex(M, L, border, "SBDF3", 1e-3, F=lambda xx, tt, k=k: g[k] * xx + h[k])
        xo, _ = o.run(x0[k].copy(), 0.0, 30)
        assert rel(got[k], xo) < 1e-11
